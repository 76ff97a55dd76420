// hj_runtime.hip -- context, scratch arena, timing, algorithm registry and
// the hj_join entry point of libhjmi.so.
//
// The registry mirrors the reference's algos[] table (src/main.c:66-143):
// a name and a function.  The reference's radix "bw_reg" twins only limit
// how many CPU threads write during partitioning to cap Optane write
// bandwidth (src/algo/phj_bw_reg.c:255-283) and give identical results; here
// they run the same GPU path as their plain names.

#include "hj_internal.h"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>

int hj_fail(hj_ctx_t *ctx, int code, const char *fmt, ...) {
	if (ctx) {
		char buf[512];
		va_list ap;
		va_start(ap, fmt);
		vsnprintf(buf, sizeof buf, fmt, ap);
		va_end(ap);
		ctx->last_error = buf;
	}
	return code;
}

int hj_reserve(hj_ctx_t *ctx, hj_buf &b, size_t bytes) {
	if (bytes == 0) bytes = 256;
	if (b.cap >= bytes) return HJ_OK;
	if (b.p) {
		hipError_t e = hipStreamSynchronize(ctx->stream);
		if (e != hipSuccess) return hj_fail(ctx, HJ_EHIP, "sync before realloc: %s", hipGetErrorString(e));
		(void)hipFree(b.p);
		b.p = nullptr;
		b.cap = 0;
	}
	size_t want = (bytes + 4095) & ~(size_t)4095;
	hipError_t e = hipMalloc(&b.p, want);
	if (e != hipSuccess) {
		(void)hipGetLastError();
		b.p = nullptr;
		return hj_fail(ctx, HJ_ENOMEM, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
	}
	b.cap = want;
	return HJ_OK;
}

int hj_scratch(hj_ctx_t *ctx, hj_buf &b, size_t bytes, void **p) {
	if (!ctx->fixed) {
		int rc = hj_reserve(ctx, b, bytes);
		*p = b.p;
		return rc;
	}
	const size_t want = (bytes + 255) & ~(size_t)255;
	if (ctx->fixed_off > ctx->fixed_cap || want > ctx->fixed_cap - ctx->fixed_off)
		return hj_fail(ctx, HJ_EINVAL, "no-allocation join needs %zu more scratch bytes than its %zu-byte region holds",
		               want - (ctx->fixed_cap - std::min(ctx->fixed_off, ctx->fixed_cap)), ctx->fixed_cap);
	*p = ctx->fixed + ctx->fixed_off;
	ctx->fixed_off += want;
	return HJ_OK;
}

void hj_kt_begin(hj_ctx_t *ctx, uint32_t kid) {
	if (!ctx->kernel_times || ctx->n_kt >= HJ_MAX_KTIMES) return;
	const int n = ctx->n_kt;
	ctx->kt_id[n] = kid;
	ctx->kt_s[n] = 4 + 2 * n;
	ctx->kt_open = true;
	(void)hipEventRecord(ctx->ev[ctx->kt_s[n]], ctx->stream);
}

void hj_kt_end(hj_ctx_t *ctx) {
	if (!ctx->kernel_times || !ctx->kt_open || ctx->n_kt >= HJ_MAX_KTIMES) return;
	const int n = ctx->n_kt;
	ctx->kt_e[n] = 4 + 2 * n + 1;
	(void)hipEventRecord(ctx->ev[ctx->kt_e[n]], ctx->stream);
	ctx->kt_open = false;
	ctx->n_kt++;
}

void hj_kt_next(hj_ctx_t *ctx, uint32_t kid) {
	if (!ctx->kernel_times || !ctx->kt_open) return;
	hj_kt_end(ctx);
	const int n = ctx->n_kt;
	if (n == 0 || n >= HJ_MAX_KTIMES) return;
	ctx->kt_id[n] = kid;
	ctx->kt_s[n] = ctx->kt_e[n - 1];
	ctx->kt_open = true;
}

void hj_kt_open_at(hj_ctx_t *ctx, uint32_t kid, int ev_index) {
	if (!ctx->kernel_times || ctx->n_kt >= HJ_MAX_KTIMES) return;
	ctx->kt_id[ctx->n_kt] = kid;
	ctx->kt_s[ctx->n_kt] = ev_index;
	ctx->kt_open = true;
}

void hj_kt_close_at(hj_ctx_t *ctx, int ev_index) {
	if (!ctx->kernel_times || !ctx->kt_open || ctx->n_kt >= HJ_MAX_KTIMES) return;
	ctx->kt_e[ctx->n_kt] = ev_index;
	ctx->kt_open = false;
	ctx->n_kt++;
}

double hj_kt_ms(hj_ctx_t *ctx, int i) { return hj_event_seconds(ctx, ctx->kt_s[i], ctx->kt_e[i]) * 1e3; }

double hj_event_seconds(hj_ctx_t *ctx, int a, int b) {
	float ms = 0.f;
	if (hipEventElapsedTime(&ms, ctx->ev[a], ctx->ev[b]) != hipSuccess) {
		(void)hipGetLastError();
		return 0.0;
	}
	return ms * 1e-3;
}

static const char *k_names[HJ_K_COUNT] = {
    "fill", "lp_build", "lp_probe", "lp_fixup", "sc_build", "sc_probe",
    "radix_hist", "radix_scan", "radix_scatter", "radix_plan", "radix_join", "radix_sortchunk",
    "radix_tilesort", "radix_gathersort", "ptr_keyrid", "ptr_retrieve",
};

extern "C" {

const char *hj_kernel_name(uint32_t id) { return id < HJ_K_COUNT ? k_names[id] : nullptr; }

const char *hj_strerror(int status) {
	switch (status) {
	case HJ_OK: return "ok";
	case HJ_EINVAL: return "invalid argument";
	case HJ_ENOMEM: return "out of memory";
	case HJ_EHIP: return "HIP runtime error";
	case HJ_EREJECT: return "request rejected by admission control";
	case HJ_ENODEV: return "no usable GPU";
	case HJ_EBUSY: return "request still running";
	default: return "unknown status";
	}
}

const char *hj_last_error(hj_ctx_t *ctx) { return ctx ? ctx->last_error.c_str() : ""; }

int hj_ctx_create(int device, hj_ctx_t **out) {
	if (!out) return HJ_EINVAL;
	*out = nullptr;
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
		(void)hipGetLastError();
		return HJ_ENODEV;
	}
	if (device < 0 || device >= n) return HJ_EINVAL;
	hj_ctx_t *ctx = new hj_ctx_t();
	ctx->device = device;
	if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&ctx->prop, device) != hipSuccess ||
	    hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
		(void)hipGetLastError();
		delete ctx;
		return HJ_EHIP;
	}
	ctx->num_cu = ctx->prop.multiProcessorCount;
	memset(ctx->tune_bits, 0xff, sizeof ctx->tune_bits);
	memset(ctx->tune_passes, 0xff, sizeof ctx->tune_passes);
	memset(ctx->tune_pass1, 0, sizeof ctx->tune_pass1);
	memset(ctx->tune_bits1, 0xff, sizeof ctx->tune_bits1);
	memset(ctx->tune_var1, 0, sizeof ctx->tune_var1);
	memset(ctx->tune_gbits, 0xff, sizeof ctx->tune_gbits);
	memset(ctx->tune_bits_z, 0xff, sizeof ctx->tune_bits_z);
	memset(ctx->tune_passes_z, 0xff, sizeof ctx->tune_passes_z);
	memset(ctx->tune_pass1_z, 0, sizeof ctx->tune_pass1_z);
	memset(ctx->tune_pscale_z, 0, sizeof ctx->tune_pscale_z);
	for (int i = 0; i < hj_ctx::NEV; i++) {
		if (hipEventCreate(&ctx->ev[i]) != hipSuccess) {
			(void)hipGetLastError();
			hj_ctx_destroy(ctx);
			return HJ_EHIP;
		}
	}
	// pinned, so the end-of-join counter read is one DMA into host memory
	// (a pageable destination is staged through the runtime's bounce buffer)
	if (hipHostMalloc(&ctx->hstats, sizeof(hj_stats), hipHostMallocDefault) != hipSuccess) {
		(void)hipGetLastError();
		ctx->hstats = nullptr;
		hj_ctx_destroy(ctx);
		return HJ_ENOMEM;
	}
	if (hj_reserve(ctx, ctx->stats, sizeof(hj_stats)) != HJ_OK) {
		hj_ctx_destroy(ctx);
		return HJ_ENOMEM;
	}
	// a failure here only disables hj_join_noalloc (it reports the error)
	(void)hj_noalloc_setup(ctx);
	*out = ctx;
	return HJ_OK;
}

int hj_ctx_destroy(hj_ctx_t *ctx) {
	if (!ctx) return HJ_EINVAL;
	(void)hipSetDevice(ctx->device);
	if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
	hj_buf *bufs[] = {&ctx->table, &ctx->pool, &ctx->aux, &ctx->part[0], &ctx->part[1], &ctx->hist,
	                  &ctx->meta, &ctx->scan_tmp, &ctx->stats, &ctx->hostcopy[0], &ctx->hostcopy[1],
	                  &ctx->pairs, &ctx->keyrid[0], &ctx->keyrid[1], &ctx->tpch[0], &ctx->tpch[1], &ctx->tpch[2],
	                  &ctx->tpch[3], &ctx->tpch[4]};
	for (hj_buf *b : bufs)
		if (b->p) (void)hipFree(b->p);
	for (int i = 0; i < hj_ctx::NEV; i++)
		if (ctx->ev[i]) (void)hipEventDestroy(ctx->ev[i]);
	if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
	if (ctx->hstats) (void)hipHostFree(ctx->hstats);
	delete ctx;
	return HJ_OK;
}

// radix joins by per-partition table (src/algo/phj.c:2515-2535)
#define HJ_RADIX_FN(NAME, TK)                                                                           \
	static int NAME(hj_ctx_t *c, const hj_relation_t *R, const hj_relation_t *S, const hj_params_t *p, \
	                hj_result_t *o) {                                                                   \
		return hj_run_radix(c, R, S, p, TK, o);                                                         \
	}
HJ_RADIX_FN(run_radix_bc, 0)
HJ_RADIX_FN(run_radix_lp, 1)
HJ_RADIX_FN(run_radix_sc, 2)
HJ_RADIX_FN(run_radix_hr, 3)
HJ_RADIX_FN(run_radix_hro, 4)
#undef HJ_RADIX_FN

// asymmetric two-pass radix joins (src/algo/phj_asymm_radix.c,
// phj_asymm_radix_bw_reg.c): NUM_RADIX_BITS_PASS1 bits in the first pass,
// the rest in the second.  The reference's default first pass is 11 bits,
// lowered to NUM_RADIX_BITS when that is smaller (phj_asymm_radix.h:10-17),
// which leaves no second pass; here that case runs one pass of the same
// bits.  The first pass is capped at 12 bits (one LDS sort tile per pass).
static int run_asymm(hj_ctx_t *c, const hj_relation_t *R, const hj_relation_t *S, const hj_params_t *p, int tk,
                     hj_result_t *o) {
	hj_params_t q = *p;
	if (q.radix_bits == 0) {
		uint32_t passes;
		hj_radix_config(c, R->tuple_num, &q.radix_bits, &passes);
	}
	uint32_t b1 = q.radix_bits_pass1 ? q.radix_bits_pass1 : 11;
	if (b1 >= q.radix_bits) {
		q.passes = q.radix_bits ? 1 : 0;
		q.radix_bits_pass1 = 0;
	} else {
		q.passes = 2;
		q.radix_bits_pass1 = std::min(b1, 12u);
	}
	return hj_run_radix(c, R, S, &q, tk, o);
}
#define HJ_ASYMM_FN(NAME, TK)                                                                           \
	static int NAME(hj_ctx_t *c, const hj_relation_t *R, const hj_relation_t *S, const hj_params_t *p, \
	                hj_result_t *o) {                                                                   \
		return run_asymm(c, R, S, p, TK, o);                                                            \
	}
HJ_ASYMM_FN(asymm_bc, 0)
HJ_ASYMM_FN(asymm_sc, 2)
HJ_ASYMM_FN(asymm_hr, 3)
HJ_ASYMM_FN(asymm_hro, 4)
#undef HJ_ASYMM_FN

// late-materialising joins: the same engines on {key, row_id} heads
// (src/algo/ptr_nphj.c, ptr_phj.c:3100-3115, ptr_phj_bw_reg.c)
#define HJ_PTR_FN(NAME, INNER)                                                                          \
	static int NAME(hj_ctx_t *c, const hj_relation_t *R, const hj_relation_t *S, const hj_params_t *p, \
	                hj_result_t *o) {                                                                   \
		return hj_run_ptr(c, R, S, p, INNER, o);                                                        \
	}
HJ_PTR_FN(ptr_nphj_sc, hj_run_nphj_sc)
HJ_PTR_FN(ptr_nphj_lp, hj_run_nphj_lp)
HJ_PTR_FN(ptr_radix_bc, run_radix_bc)
HJ_PTR_FN(ptr_radix_lp, run_radix_lp)
HJ_PTR_FN(ptr_radix_sc, run_radix_sc)
HJ_PTR_FN(ptr_radix_hr, run_radix_hr)
#undef HJ_PTR_FN

struct hj_algo_entry {
	const char *name;
	hj_algo_fn fn;
};

static const hj_algo_entry k_algos[] = {
    {"nphj_sc", hj_run_nphj_sc},
    {"nphj_lp", hj_run_nphj_lp},
    {"phj_radix_bc", run_radix_bc},
    {"phj_radix_lp", run_radix_lp},
    {"phj_radix_sc", run_radix_sc},
    {"phj_radix_hr", run_radix_hr},
    {"phj_radix_hro", run_radix_hro},
    {"phj_radix_bw_reg_bc", run_radix_bc},
    {"phj_radix_bw_reg_lp", run_radix_lp},
    {"phj_radix_bw_reg_sc", run_radix_sc},
    {"phj_radix_bw_reg_hr", run_radix_hr},
    {"phj_radix_bw_reg_hro", run_radix_hro},
    {"phj_asymm_radix_bc", asymm_bc},
    {"phj_asymm_radix_sc", asymm_sc},
    {"phj_asymm_radix_hr", asymm_hr},
    {"phj_asymm_radix_hro", asymm_hro},
    {"phj_asymm_radix_bw_reg_bc", asymm_bc},
    {"phj_asymm_radix_bw_reg_sc", asymm_sc},
    {"phj_asymm_radix_bw_reg_hr", asymm_hr},
    {"phj_asymm_radix_bw_reg_hro", asymm_hro},
    {"ptr_nphj_sc", ptr_nphj_sc},
    {"ptr_nphj_lp", ptr_nphj_lp},
    {"ptr_phj_radix_bc", ptr_radix_bc},
    {"ptr_phj_radix_sc", ptr_radix_sc},
    {"ptr_phj_radix_lp", ptr_radix_lp},
    {"ptr_phj_radix_hr", ptr_radix_hr},
    {"ptr_phj_radix_bw_reg_bc", ptr_radix_bc},
    {"ptr_phj_radix_bw_reg_sc", ptr_radix_sc},
    {"ptr_phj_radix_bw_reg_lp", ptr_radix_lp},
    {"ptr_phj_radix_bw_reg_hr", ptr_radix_hr},
    {nullptr, nullptr},
};

int hj_algo_count(void) { return (int)(sizeof k_algos / sizeof k_algos[0]) - 1; }

const char *hj_algo_name(int idx) {
	if (idx < 0 || idx >= hj_algo_count()) return nullptr;
	return k_algos[idx].name;
}

bool hj_algo_is_lp(const char *algo, bool *radix) {
	*radix = strncmp(algo, "phj_radix_", 10) == 0;
	bool known = false;
	for (const hj_algo_entry *a = k_algos; a->name; a++)
		if (strcmp(a->name, algo) == 0) known = true;
	if (!known) *radix = false;
	const size_t n = strlen(algo);
	return n >= 3 && strcmp(algo + n - 3, "_lp") == 0;
}

// the reference's payload sweep: 16..512-byte tuples (logs/payload-logs)
static bool width_ok(uint32_t w) { return w == 16 || w == 32 || w == 64 || w == 128 || w == 256 || w == 512; }

int hj_join(hj_ctx_t *ctx, const char *algo, const hj_relation_t *R, const hj_relation_t *S,
            const hj_params_t *params, hj_result_t *out) {
	if (!ctx || !algo || !R || !S || !out) return hj_fail(ctx, HJ_EINVAL, "null argument");
	std::lock_guard<std::mutex> lk(ctx->mu);
	const hj_algo_entry *e = nullptr;
	for (const hj_algo_entry *a = k_algos; a->name; a++)
		if (strcmp(a->name, algo) == 0) e = a;
	if (!e) return hj_fail(ctx, HJ_EINVAL, "unknown algo '%s'", algo);
	if (!width_ok(R->tuple_size) || R->tuple_size != S->tuple_size)
		return hj_fail(ctx, HJ_EINVAL, "tuple sizes %u/%u unsupported (16..512, equal)", R->tuple_size,
		               S->tuple_size);
	if ((R->tuple_num && !R->tuples) || (S->tuple_num && !S->tuples))
		return hj_fail(ctx, HJ_EINVAL, "null tuples pointer");
	if (R->location > HJ_LOC_HOST || S->location > HJ_LOC_HOST) return hj_fail(ctx, HJ_EINVAL, "bad location");
	hj_params_t prm{};
	if (params) prm = *params;
	if (prm.passes > 2) return hj_fail(ctx, HJ_EINVAL, "passes must be 0, 1 or 2");
	if (prm.radix_bits > 20) return hj_fail(ctx, HJ_EINVAL, "radix_bits must be <= 20");
	memset(out, 0, sizeof *out);
	HJ_HIP(ctx, hipSetDevice(ctx->device));

	// Host-resident relations are copied in first; the copy is outside the
	// timed phases, as the reference's load is (src/algo/phj.c:1885-1915).
	hj_relation_t rel[2] = {*R, *S};
	for (int i = 0; i < 2; i++) {
		if (rel[i].location != HJ_LOC_HOST || rel[i].tuple_num == 0) continue;
		size_t bytes = (size_t)rel[i].tuple_num * rel[i].tuple_size;
		int rc = hj_reserve(ctx, ctx->hostcopy[i], bytes);
		if (rc) return rc;
		HJ_HIP(ctx, hipMemcpyAsync(ctx->hostcopy[i].p, rel[i].tuples, bytes, hipMemcpyHostToDevice, ctx->stream));
		rel[i].tuples = ctx->hostcopy[i].p;
		rel[i].location = HJ_LOC_DEVICE;
	}
	ctx->kernel_times = (prm.flags & HJ_F_KERNEL_TIMES) != 0;
	ctx->n_kt = 0;
	ctx->kt_open = false;
	int rc = e->fn(ctx, &rel[0], &rel[1], &prm, out);
	if (rc == HJ_OK) {
		out->n_ktimes = (uint32_t)ctx->n_kt;
		for (int i = 0; i < ctx->n_kt; i++) {
			out->ktime_id[i] = ctx->kt_id[i];
			out->ktime_ms[i] = hj_kt_ms(ctx, i);
		}
	}
	ctx->kernel_times = false;
	return rc;
}

int hj_dev_alloc(hj_ctx_t *ctx, uint64_t bytes, void **ptr) {
	if (!ctx || !ptr) return HJ_EINVAL;
	HJ_HIP(ctx, hipSetDevice(ctx->device));
	hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
	if (e != hipSuccess) {
		(void)hipGetLastError();
		*ptr = nullptr;
		return hj_fail(ctx, HJ_ENOMEM, "hipMalloc(%llu): %s", (unsigned long long)bytes, hipGetErrorString(e));
	}
	return HJ_OK;
}

int hj_dev_free(hj_ctx_t *ctx, void *ptr) {
	if (!ctx) return HJ_EINVAL;
	if (!ptr) return HJ_OK;
	HJ_HIP(ctx, hipSetDevice(ctx->device));
	HJ_HIP(ctx, hipStreamSynchronize(ctx->stream));
	HJ_HIP(ctx, hipFree(ptr));
	return HJ_OK;
}

int hj_copy_h2d(hj_ctx_t *ctx, void *dst, const void *src, uint64_t bytes) {
	if (!ctx || (bytes && (!dst || !src))) return HJ_EINVAL;
	if (!bytes) return HJ_OK;
	HJ_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
	HJ_HIP(ctx, hipStreamSynchronize(ctx->stream));
	return HJ_OK;
}

int hj_copy_d2h(hj_ctx_t *ctx, void *dst, const void *src, uint64_t bytes) {
	if (!ctx || (bytes && (!dst || !src))) return HJ_EINVAL;
	if (!bytes) return HJ_OK;
	HJ_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
	HJ_HIP(ctx, hipStreamSynchronize(ctx->stream));
	return HJ_OK;
}

int hj_sync(hj_ctx_t *ctx) {
	if (!ctx) return HJ_EINVAL;
	HJ_HIP(ctx, hipStreamSynchronize(ctx->stream));
	return HJ_OK;
}

int hj_generate(hj_ctx_t *ctx, const hj_gen_spec_t *spec, void *r_dev, void *s_dev) {
	if (!ctx || !spec) return HJ_EINVAL;
	if (!width_ok(spec->width)) return hj_fail(ctx, HJ_EINVAL, "width %u unsupported", spec->width);
	if ((spec->r_num && !r_dev) || (spec->s_num && !s_dev)) return hj_fail(ctx, HJ_EINVAL, "null buffer");
	if (spec->workload > HJ_WL_RANDOM) return hj_fail(ctx, HJ_EINVAL, "unknown workload %u", spec->workload);
	if (spec->workload == HJ_WL_RANDOM && spec->key_range == 0)
		return hj_fail(ctx, HJ_EINVAL, "random workload needs key_range >= 1");
	if (spec->workload != HJ_WL_RANDOM && spec->s_num && spec->r_num == 0)
		return hj_fail(ctx, HJ_EINVAL, "S keys are drawn from R: |R| must be >= 1");
	std::lock_guard<std::mutex> lk(ctx->mu);
	HJ_HIP(ctx, hipSetDevice(ctx->device));
	int rc = hj_launch_generate(ctx, spec, r_dev, s_dev);
	if (rc) return rc;
	HJ_HIP(ctx, hipStreamSynchronize(ctx->stream));
	return HJ_OK;
}

int hj_device_info(hj_ctx_t *ctx, hj_device_info_t *info) {
	if (!ctx || !info) return HJ_EINVAL;
	memset(info, 0, sizeof *info);
	info->cu_count = ctx->prop.multiProcessorCount;
	info->lds_per_block = (int)ctx->prop.sharedMemPerBlock;
	info->lds_per_cu = (int)ctx->prop.maxSharedMemoryPerMultiProcessor;
	info->warp_size = ctx->prop.warpSize;
	info->hbm_bytes = ctx->prop.totalGlobalMem;
	info->clock_khz = ctx->prop.clockRate;
	snprintf(info->arch, sizeof info->arch, "%s", ctx->prop.gcnArchName);
	return HJ_OK;
}

}  // extern "C"
