// lela_amd/csrc/ctx.hip -- context lifetime, device workspace, staging,
// profiling events and the error channel of the C ABI.
#include "common.hpp"

#include <cstdarg>

namespace mlz {

int set_error(mlz_ctx *ctx, int code, const char *fmt, ...)
{
	char buf[512];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof buf, fmt, ap);
	va_end(ap);
	if (ctx)
		ctx->last_error = buf;
	return code;
}

int check_launch(mlz_ctx *ctx, const char *what)
{
	hipError_t e = hipGetLastError();
	if (e != hipSuccess)
		return set_error(ctx, MLZ_E_HIP, "launch of %s failed: %s", what, hipGetErrorString(e));
	return MLZ_OK;
}

int ws_get(mlz_ctx *ctx, int slot, size_t bytes, void **out)
{
	if (bytes == 0)
		bytes = 256;
	if (ctx->ws_bytes[slot] < bytes) {
		if (ctx->ws[slot]) {
			// The old buffer may still be read by queued work.
			MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
			MLZ_HIP(ctx, hipFree(ctx->ws[slot]));
			ctx->ws[slot] = nullptr;
			ctx->ws_bytes[slot] = 0;
		}
		size_t sz = round_up(bytes, 1 << 20);
		void *p = nullptr;
		hipError_t e = hipMalloc(&p, sz);
		if (e != hipSuccess) {
			(void) hipGetLastError();
			return set_error(ctx, MLZ_E_NOMEM, "hipMalloc(%zu) for workspace slot %d: %s", sz,
					 slot, hipGetErrorString(e));
		}
		ctx->ws[slot] = p;
		ctx->ws_bytes[slot] = sz;
	}
	*out = ctx->ws[slot];
	return MLZ_OK;
}

static hipEvent_t get_event(mlz_ctx *ctx)
{
	if (!ctx->free_events.empty()) {
		hipEvent_t e = ctx->free_events.back();
		ctx->free_events.pop_back();
		return e;
	}
	hipEvent_t e = nullptr;
	if (hipEventCreate(&e) != hipSuccess)
		return nullptr;
	return e;
}

void prof_begin(mlz_ctx *ctx, const char *name, PendingEv *ev)
{
	ev->name = name;
	ev->e0 = ev->e1 = nullptr;
	if (!ctx->prof_on || ctx->noalloc || !name)
		return;
	ev->e0 = get_event(ctx);
	ev->e1 = get_event(ctx);
	if (ev->e0)
		(void) hipEventRecord(ev->e0, ctx->stream);
}

void prof_end(mlz_ctx *ctx, PendingEv *ev)
{
	if (!ctx->prof_on || ctx->noalloc || !ev->e0 || !ev->e1)
		return;
	(void) hipEventRecord(ev->e1, ctx->stream);
	ctx->pending.push_back(*ev);
}

static int prof_flush(mlz_ctx *ctx)
{
	if (ctx->pending.empty())
		return MLZ_OK;
	MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
	for (auto &pe : ctx->pending) {
		float ms = 0.f;
		(void) hipEventElapsedTime(&ms, pe.e0, pe.e1);
		bool found = false;
		for (auto &r : ctx->prof)
			if (r.name == pe.name) {
				r.count++;
				r.ms += ms;
				found = true;
				break;
			}
		if (!found)
			ctx->prof.push_back(ProfRec{pe.name, 1, (double) ms});
		ctx->free_events.push_back(pe.e0);
		ctx->free_events.push_back(pe.e1);
	}
	ctx->pending.clear();
	return MLZ_OK;
}

int stage_in(mlz_ctx *ctx, int slot, const void *host, size_t rows, size_t cols, size_t ld,
	     size_t esize, void **dev)
{
	MLZ_TRY(ws_get(ctx, slot, rows * cols * esize, dev));
	if (rows && cols)
		MLZ_HIP(ctx, hipMemcpy2DAsync(*dev, cols * esize, host, ld * esize, cols * esize, rows,
					      hipMemcpyHostToDevice, ctx->stream));
	return MLZ_OK;
}

int stage_out(mlz_ctx *ctx, void *host, const void *dev, size_t rows, size_t cols, size_t ld,
	      size_t esize)
{
	if (rows && cols)
		MLZ_HIP(ctx, hipMemcpy2DAsync(host, ld * esize, dev, cols * esize, cols * esize, rows,
					      hipMemcpyDeviceToHost, ctx->stream));
	MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
	return MLZ_OK;
}

int aux_stream(mlz_ctx *ctx, hipStream_t *out)
{
	if (!ctx->aux_stream) {
		// Default priority only: streams of different priorities are not used.
		const hipError_t e = hipStreamCreateWithFlags(&ctx->aux_stream, hipStreamNonBlocking);
		if (e != hipSuccess) {
			ctx->aux_stream = nullptr;
	ctx->copy_stream = nullptr;
			return set_error(ctx, MLZ_E_HIP, "hipStreamCreateWithFlags: %s", hipGetErrorString(e));
		}
	}
	*out = ctx->aux_stream;
	return MLZ_OK;
}

int copy_stream(mlz_ctx *ctx, hipStream_t *out)
{
	if (!ctx->copy_stream) {
		const hipError_t e = hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking);
		if (e != hipSuccess) {
			ctx->copy_stream = nullptr;
			return set_error(ctx, MLZ_E_HIP, "hipStreamCreateWithFlags: %s", hipGetErrorString(e));
		}
	}
	*out = ctx->copy_stream;
	return MLZ_OK;
}

}  // namespace mlz

using namespace mlz;

extern "C" {

const char *mlz_version(void) { return "lela_amd 0.1 (gfx950)"; }

int mlz_ctx_create(mlz_ctx **out, int device, const char *config)
{
	(void) config;
	if (!out)
		return MLZ_E_ARG;
	*out = nullptr;
	int ndev = 0;
	if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
		(void) hipGetLastError();
		return MLZ_E_HIP;
	}
	if (device < 0 || device >= ndev)
		return MLZ_E_ARG;
	if (hipSetDevice(device) != hipSuccess)
		return MLZ_E_HIP;
	mlz_ctx *ctx = new mlz_ctx();
	ctx->device = device;
	ctx->ptr_mode = MLZ_PTR_HOST;
	ctx->prof_on = false;
	for (int i = 0; i < WS_NUM_SLOTS; ++i) {
		ctx->ws[i] = nullptr;
		ctx->ws_bytes[i] = 0;
	}
	ctx->host_pinned = nullptr;
	ctx->host_pinned_bytes = 0;
	ctx->i8_variant = 0;
	ctx->upd_nblk = 1;
	ctx->upd_lds = -1;
	ctx->rref_fast_threads = 1024;
	ctx->invtab_p = 0;
	ctx->gf2_kernel = 0;
	ctx->tuner = nullptr;
	ctx->noalloc = false;
	ctx->fixed_ws = nullptr;
	ctx->fixed_ws_bytes = 0;
	ctx->aux_stream = nullptr;
	// Default priority only: streams of different priorities are not used.
	if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
		delete ctx;
		return MLZ_E_HIP;
	}
	if (hipEventCreate(&ctx->timer0) != hipSuccess || hipEventCreate(&ctx->timer1) != hipSuccess) {
		delete ctx;
		return MLZ_E_HIP;
	}
	hipDeviceProp_t prop;
	ctx->num_cus = 256;
	if (hipGetDeviceProperties(&prop, device) == hipSuccess)
		ctx->num_cus = prop.multiProcessorCount;
	// A failed lookup leaves fixed_ws null: only the no-allocation entry
	// point needs it, and it reports the failure (the message is kept).
	(void) noalloc_scratch_init(ctx);
	*out = ctx;
	return MLZ_OK;
}

int mlz_ctx_destroy(mlz_ctx *ctx)
{
	if (!ctx)
		return MLZ_E_ARG;
	(void) hipSetDevice(ctx->device);
	(void) hipStreamSynchronize(ctx->stream);
	tuner_free(ctx);
	for (int i = 0; i < WS_NUM_SLOTS; ++i)
		if (ctx->ws[i])
			(void) hipFree(ctx->ws[i]);
	for (auto &pe : ctx->pending) {
		(void) hipEventDestroy(pe.e0);
		(void) hipEventDestroy(pe.e1);
	}
	for (auto e : ctx->free_events)
		(void) hipEventDestroy(e);
	(void) hipEventDestroy(ctx->timer0);
	(void) hipEventDestroy(ctx->timer1);
	if (ctx->host_pinned)
		(void) hipHostFree(ctx->host_pinned);
	if (ctx->aux_stream)
		(void) hipStreamDestroy(ctx->aux_stream);
	if (ctx->copy_stream)
		(void) hipStreamDestroy(ctx->copy_stream);
	(void) hipStreamDestroy(ctx->stream);
	delete ctx;
	return MLZ_OK;
}

int mlz_ctx_set_ptr_mode(mlz_ctx *ctx, int mode)
{
	if (!ctx || (mode != MLZ_PTR_HOST && mode != MLZ_PTR_DEVICE))
		return MLZ_E_ARG;
	ctx->ptr_mode = mode;
	return MLZ_OK;
}

int mlz_ctx_set_option(mlz_ctx *ctx, const char *key, int64_t value)
{
	if (!ctx || !key)
		return MLZ_E_ARG;
	if (strcmp(key, "gemm.i8.variant") == 0) {
		if (value < 0 || value > 11)
			return set_error(ctx, MLZ_E_ARG, "gemm.i8.variant must be 0..11");
		ctx->i8_variant = (int) value;
		return MLZ_OK;
	}
	if (strcmp(key, "rref.search.threads") == 0) {
		if (value != 256 && value != 512 && value != 1024)
			return set_error(ctx, MLZ_E_ARG, "rref.search.threads must be 256, 512 or 1024");
		ctx->rref_fast_threads = (int) value;
		return MLZ_OK;
	}
	if (strcmp(key, "rref.update.blocks") == 0) {
		if (value != 1 && value != 2 && value != 4)
			return set_error(ctx, MLZ_E_ARG, "rref.update.blocks must be 1, 2 or 4");
		ctx->upd_nblk = (int) value;
		return MLZ_OK;
	}
	if (strcmp(key, "update.i8.lds") == 0) {
		if (value != -1 && value != 0 && value != 1 && value != 2 && value != 4)
			return set_error(ctx, MLZ_E_ARG, "update.i8.lds must be -1, 0, 1, 2 or 4");
		ctx->upd_lds = (int) value;
		return MLZ_OK;
	}
	if (strcmp(key, "gf2.gemm.kernel") == 0) {
		if (value != 0 && value != 1)
			return set_error(ctx, MLZ_E_ARG, "gf2.gemm.kernel must be 0 or 1");
		ctx->gf2_kernel = (int) value;
		return MLZ_OK;
	}
	return set_error(ctx, MLZ_E_ARG, "unknown option '%s'", key);
}

int mlz_ctx_synchronize(mlz_ctx *ctx)
{
	if (!ctx)
		return MLZ_E_ARG;
	MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
	return MLZ_OK;
}

const char *mlz_last_error(mlz_ctx *ctx)
{
	return ctx ? ctx->last_error.c_str() : "null context";
}

int mlz_dev_alloc(mlz_ctx *ctx, size_t bytes, void **ptr)
{
	if (!ctx || !ptr)
		return MLZ_E_ARG;
	hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
	if (e != hipSuccess) {
		(void) hipGetLastError();
		return set_error(ctx, MLZ_E_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
	}
	return MLZ_OK;
}

int mlz_dev_free(mlz_ctx *ctx, void *ptr)
{
	if (!ctx)
		return MLZ_E_ARG;
	MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
	MLZ_HIP(ctx, hipFree(ptr));
	return MLZ_OK;
}

int mlz_memcpy_htod(mlz_ctx *ctx, void *dst, const void *src, size_t bytes)
{
	if (!ctx)
		return MLZ_E_ARG;
	MLZ_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
	MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
	return MLZ_OK;
}

int mlz_memcpy_dtoh(mlz_ctx *ctx, void *dst, const void *src, size_t bytes)
{
	if (!ctx)
		return MLZ_E_ARG;
	MLZ_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
	MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
	return MLZ_OK;
}

int mlz_memcpy_dtod(mlz_ctx *ctx, void *dst, const void *src, size_t bytes)
{
	if (!ctx)
		return MLZ_E_ARG;
	MLZ_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
	MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
	return MLZ_OK;
}

int mlz_ctx_trim(mlz_ctx *ctx)
{
	if (!ctx)
		return MLZ_E_ARG;
	MLZ_HIP(ctx, hipStreamSynchronize(ctx->stream));
	for (int i = 0; i < WS_NUM_SLOTS; ++i)
		if (ctx->ws[i]) {
			MLZ_HIP(ctx, hipFree(ctx->ws[i]));
			ctx->ws[i] = nullptr;
			ctx->ws_bytes[i] = 0;
		}
	return MLZ_OK;
}

int mlz_prof_enable(mlz_ctx *ctx, int on)
{
	if (!ctx)
		return MLZ_E_ARG;
	ctx->prof_on = on != 0;
	return MLZ_OK;
}

int mlz_prof_reset(mlz_ctx *ctx)
{
	if (!ctx)
		return MLZ_E_ARG;
	MLZ_TRY(prof_flush(ctx));
	ctx->prof.clear();
	return MLZ_OK;
}

int mlz_prof_query(mlz_ctx *ctx, const char *name, uint64_t *count, double *ms)
{
	if (!ctx || !name)
		return MLZ_E_ARG;
	MLZ_TRY(prof_flush(ctx));
	uint64_t c = 0;
	double t = 0;
	for (auto &r : ctx->prof)
		if (r.name.find(name) != std::string::npos) {
			c += r.count;
			t += r.ms;
		}
	if (count)
		*count = c;
	if (ms)
		*ms = t;
	return MLZ_OK;
}

int mlz_timer_start(mlz_ctx *ctx)
{
	if (!ctx)
		return MLZ_E_ARG;
	MLZ_HIP(ctx, hipEventRecord(ctx->timer0, ctx->stream));
	return MLZ_OK;
}

int mlz_timer_stop(mlz_ctx *ctx, double *ms)
{
	if (!ctx)
		return MLZ_E_ARG;
	MLZ_HIP(ctx, hipEventRecord(ctx->timer1, ctx->stream));
	MLZ_HIP(ctx, hipEventSynchronize(ctx->timer1));
	float t = 0.f;
	MLZ_HIP(ctx, hipEventElapsedTime(&t, ctx->timer0, ctx->timer1));
	if (ms)
		*ms = t;
	return MLZ_OK;
}

}  // extern "C"
