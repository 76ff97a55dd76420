// hj_noalloc.hip -- hj_join_noalloc: a non-partitioned join (nphj_sc,
// nphj_lp) that allocates nothing.  For callers that cannot call hipMalloc
// while they run: code under HIP graph capture, tight serving loops,
// harnesses that forbid allocation.
//
// Its scratch -- the join's counters, hash table, overflow pool and
// duplicate fix-up indices -- lives in one file-scope device array,
// `scratch`, of HJ_NOALLOC_SCRATCH_BYTES.  Each GPU the library runs on holds
// its own copy until the process ends.  In host code `scratch` names a
// host-side stand-in, not a device address, so the device address is
// resolved once per device with hipGetSymbolAddress when a context is created
// (hj_noalloc_setup, with the context's device current) and only that address
// is handed to kernels and copies.  Every call on a GPU shares that GPU's
// array: calls that overlap on one GPU (two contexts or host threads at once)
// race.  The array keeps the last call's contents; the runners clear what must
// start at zero (counters, table) with fill kernels on every call.

#include "hj_internal.h"

#include <string.h>

#include <mutex>

static __device__ __align__(256) uint8_t scratch[HJ_NOALLOC_SCRATCH_BYTES];
static_assert(sizeof(hj_stats) <= 256, "the counters take the first 256 bytes of the region");

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};
hipError_t g_err[kMaxDevices] = {};
bool g_done[kMaxDevices] = {};
}  // namespace

int hj_noalloc_setup(hj_ctx_t *ctx) {
	if (ctx->device < 0 || ctx->device >= kMaxDevices) return HJ_EINVAL;
	std::lock_guard<std::mutex> lk(g_mu);
	if (g_done[ctx->device]) return g_err[ctx->device] == hipSuccess ? HJ_OK : HJ_EHIP;
	void *p = nullptr;
	const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
	if (e != hipSuccess) {
		(void)hipGetLastError();
		p = nullptr;
	}
	g_ptr[ctx->device] = p;
	g_err[ctx->device] = e;
	g_done[ctx->device] = true;
	return e == hipSuccess ? HJ_OK : HJ_EHIP;
}

extern "C" int hj_join_noalloc(hj_ctx_t *ctx, const char *algo, const hj_relation_t *R, const hj_relation_t *S,
                               const hj_params_t *params, hj_result_t *out) {
	if (!ctx || !algo || !R || !S || !out) return hj_fail(ctx, HJ_EINVAL, "null argument");
	std::lock_guard<std::mutex> lk(ctx->mu);
	hj_algo_fn fn = nullptr;
	if (strcmp(algo, "nphj_sc") == 0) fn = hj_run_nphj_sc;
	else if (strcmp(algo, "nphj_lp") == 0) fn = hj_run_nphj_lp;
	if (!fn) return hj_fail(ctx, HJ_EINVAL, "hj_join_noalloc runs nphj_sc or nphj_lp, not '%s'", algo);
	const uint32_t w = R->tuple_size;
	if (!(w == 16 || w == 32 || w == 64 || w == 128 || w == 256 || w == 512) || w != S->tuple_size)
		return hj_fail(ctx, HJ_EINVAL, "tuple sizes %u/%u unsupported (16..512, equal)", R->tuple_size,
		               S->tuple_size);
	if ((R->tuple_num && !R->tuples) || (S->tuple_num && !S->tuples))
		return hj_fail(ctx, HJ_EINVAL, "null tuples pointer");
	if (R->location != HJ_LOC_DEVICE || S->location != HJ_LOC_DEVICE)
		return hj_fail(ctx, HJ_EINVAL, "hj_join_noalloc takes device-resident relations only");
	hj_params_t prm{};
	if (params) prm = *params;
	if (prm.flags & ~(uint32_t)HJ_F_KERNEL_TIMES) return hj_fail(ctx, HJ_EINVAL, "unsupported flags 0x%x", prm.flags);
	if (prm.lp_slots_log2 > 40) return hj_fail(ctx, HJ_EINVAL, "lp_slots_log2 must be <= 40");
	// the worst case up front, so a rejected call launches nothing: LP takes
	// the slot table plus the duplicate fix-up indices, SC the bucket table
	// plus one overflow node per build tuple
	const uint64_t nr = R->tuple_num;
	const uint64_t np2 = hj_next_pow2(nr ? nr : 1);
	uint64_t need = 256;
	if (fn == hj_run_nphj_lp) {
		const uint64_t slots = prm.lp_slots_log2 ? (1ull << prm.lp_slots_log2) : np2;
		need += ((slots * 16 + 255) & ~255ull) + ((slots * 8 + 255) & ~255ull);
	} else {
		const uint64_t nb = np2 / 2 ? np2 / 2 : 1;
		need += ((nb * 64 + 255) & ~255ull) + (((nr + 1024) * 64 + 255) & ~255ull);
	}
	if (need > HJ_NOALLOC_SCRATCH_BYTES)
		return hj_fail(ctx, HJ_EINVAL,
		               "%s with |R| = %llu needs %llu scratch bytes; hj_join_noalloc holds %llu (larger joins: hj_join)",
		               algo, (unsigned long long)nr, (unsigned long long)need,
		               (unsigned long long)HJ_NOALLOC_SCRATCH_BYTES);
	void *base = nullptr;
	{
		std::lock_guard<std::mutex> g(g_mu);
		if (ctx->device >= 0 && ctx->device < kMaxDevices && g_done[ctx->device]) {
			if (g_err[ctx->device] != hipSuccess)
				return hj_fail(ctx, HJ_EHIP, "hipGetSymbolAddress(scratch) failed: %s",
				               hipGetErrorString(g_err[ctx->device]));
			base = g_ptr[ctx->device];
		}
	}
	if (!base) return hj_fail(ctx, HJ_EHIP, "no-allocation scratch not resolved on device %d", ctx->device);
	memset(out, 0, sizeof *out);
	HJ_HIP(ctx, hipSetDevice(ctx->device));
	ctx->fixed = (uint8_t *)base;
	ctx->fixed_cap = HJ_NOALLOC_SCRATCH_BYTES;
	ctx->fixed_off = 256;  // the counters (hj_stats_dev)
	ctx->kernel_times = (prm.flags & HJ_F_KERNEL_TIMES) != 0;
	ctx->n_kt = 0;
	ctx->kt_open = false;
	const int rc = fn(ctx, R, S, &prm, out);
	if (rc == HJ_OK) {
		out->n_ktimes = (uint32_t)ctx->n_kt;
		for (int i = 0; i < ctx->n_kt; i++) {
			out->ktime_id[i] = ctx->kt_id[i];
			out->ktime_ms[i] = hj_kt_ms(ctx, i);
		}
	}
	ctx->kernel_times = false;
	ctx->fixed = nullptr;
	ctx->fixed_cap = ctx->fixed_off = 0;
	return rc;
}
