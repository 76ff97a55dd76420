// hj_nphj.hip -- non-partitioned hash joins on one global table in HBM.
//
//  nphj_lp  (reference: src/algo/nphj.c:1056-1269 build/probe, :1487-1848
//           driver).  Open addressing with linear probing over
//           next_pow_2(|R|) slots, identity hash key & mask.  The reference
//           claims a slot with a CAS on a separate hashkey word and then
//           stores the 16-byte tuple (32-byte entries); here a slot is the
//           8-byte {u32 key, u32 payload} pair, claimed whole by one 64-bit
//           atomicCAS against an empty marker (all ones), when every R key
//           is below 2^32 - 1 and every payload fits 32 bits; otherwise the
//           build flags the relation and a 16-byte {key, payload} table is
//           built instead (claim: CAS of the key word), where R tuples whose
//           key equals the marker are counted on the side.
//           Probes stop at the first match, as lp_lookup_with_pos does
//           (:1162-1165); a miss adds nothing (the reference dereferences
//           NULL there, :1253).  Duplicate build keys keep one slot per key
//           holding the payload of the smallest R row index (the match a
//           sequential build would find first); they are fixed up in two
//           extra passes that run only when the build saw a duplicate.
//
//  nphj_sc  (reference: src/algo/nphj.c:132-442, :677-1037).  Separate
//           chaining.  The reference uses 2-tuple buckets with a spinlock
//           and per-thread overflow buffers; here a bucket is one 64-byte
//           line {3 tuples, u32 count, u32 next}: slots are claimed with an
//           atomicAdd on count, overflow nodes come from a pool cursor and
//           are pushed wait-free with atomicExch on next.  Probes compare
//           every occupied slot along the chain (all matches count).
//
// Both probe kernels read S with coalesced 16-byte loads, keep U tuples in
// flight per lane, and reduce count/checksum with wave shuffles + one 64-bit
// atomic per block.

#include "hj_device.h"
#include "hj_internal.h"

#define LP_EMPTY 0xFFFFFFFFFFFFFFFFULL
#define NPHJ_THREADS 256
#define PROBE_U 4

// ---------------------------------------------------------------- utilities
// Algorithmic bytes of reading the 16-byte head of a W-byte tuple: all of a
// tuple up to one 64-byte HBM burst (W + 16 per tuple for NPHJ up to W = 64;
// wider tuples are never read past the burst that holds the head).
static inline uint32_t head_bytes(uint32_t W) { return W < 64 ? W : 64; }
__global__ void k_fill_u4(uint4 *p, uint64_t n, uint4 v) {
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

__global__ void k_stats_reset(hj_stats *st) {
	unsigned long long *p = (unsigned long long *)st;
	for (unsigned i = threadIdx.x; i < sizeof(hj_stats) / 8; i += blockDim.x) p[i] = 0;
	if (threadIdx.x == 0) st->sent_min = ~0ULL;
}

static unsigned grid_for(hj_ctx_t *ctx, uint64_t n, unsigned threads, unsigned per_cu = 16) {
	uint64_t b = (n + threads - 1) / threads;
	uint64_t cap = (uint64_t)ctx->num_cu * per_cu;
	if (b > cap) b = cap;
	if (b == 0) b = 1;
	return (unsigned)b;
}

__global__ __launch_bounds__(256) void k_copy_u4(uint4 *__restrict__ dst, const uint4 *__restrict__ src, uint64_t n) {
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
		hj_store_stream(&dst[i], hj_load_stream(&src[i]));
}

extern "C" int hj_copy_d2d(hj_ctx_t *ctx, void *dst, const void *src, uint64_t bytes, double *ms) {
	if (!ctx || (bytes && (!dst || !src)) || (bytes & 15)) return HJ_EINVAL;
	if (!bytes) return HJ_OK;
	std::lock_guard<std::mutex> lk(ctx->mu);
	HJ_HIP(ctx, hipSetDevice(ctx->device));
	HJ_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
	hipLaunchKernelGGL(k_copy_u4, dim3(grid_for(ctx, bytes / 16, 256, 8)), dim3(256), 0, ctx->stream, (uint4 *)dst,
	                   (const uint4 *)src, bytes / 16);
	HJ_HIP(ctx, hipGetLastError());
	HJ_HIP(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
	HJ_HIP(ctx, hipStreamSynchronize(ctx->stream));
	if (ms) *ms = hj_event_seconds(ctx, 0, 1) * 1e3;
	return HJ_OK;
}

int hj_fill16(hj_ctx_t *ctx, void *p, uint64_t n16, uint4 v) {
	if (!n16) return HJ_OK;
	hipLaunchKernelGGL(k_fill_u4, dim3(grid_for(ctx, n16, 256, 32)), dim3(256), 0, ctx->stream, (uint4 *)p, n16, v);
	return HJ_OK;
}

int hj_stats_reset(hj_ctx_t *ctx) {
	hipLaunchKernelGGL(k_stats_reset, dim3(1), dim3(64), 0, ctx->stream, hj_stats_dev(ctx));
	return HJ_OK;
}

int hj_stats_read(hj_ctx_t *ctx, hj_stats *h) {
	HJ_HIP(ctx, hipMemcpyAsync(ctx->hstats, hj_stats_dev(ctx), sizeof *h, hipMemcpyDeviceToHost, ctx->stream));
	HJ_HIP(ctx, hipStreamSynchronize(ctx->stream));
	*h = *(const hj_stats *)ctx->hstats;
	return HJ_OK;
}

// ------------------------------------------------------------------ nphj_lp
// Home slot of a key.  nphj_lp uses the reference's identity hash key & mask;
// the exact first-match path of the radix, batch and serving lp joins (MIX)
// uses a multiplicative (Fibonacci) hash, (key * 2^64/phi) >> (64 - log2
// slots), so that keys sharing their low bits (multiples of 2^12, say) do not
// all start at a few slots: with the identity hash 30000 such keys in 2^15
// slots form 8 clusters, and the build's claims (device-scope CAS on a few hot
// lines) grow with the square of the cluster size.  The result does not
// depend on the hash: every key keeps the payload of its smallest R row.
template <bool MIX>
__device__ __forceinline__ uint64_t lp_home(uint64_t key, uint64_t mask, uint32_t shift) {
	return MIX ? (key * 0x9E3779B97F4A7C15ULL) >> shift : key & mask;
}

template <int WV, bool MIX>
__global__ __launch_bounds__(NPHJ_THREADS) void k_lp_build(const uint4 *__restrict__ R, uint64_t n,
                                                           unsigned long long *__restrict__ tab, uint64_t mask,
                                                           uint32_t shift, hj_stats *st) {
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	bool dup = false, full = false;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
		const uint4 t = R[i * WV];
		const uint64_t key = hj_key_of(t), pay = hj_pay_of(t);
		if (key == LP_EMPTY) {
			atomicAdd(&st->sent_cnt, 1ULL);
			atomicMin(&st->sent_min, (unsigned long long)i);
			continue;
		}
		uint64_t pos = lp_home<MIX>(key, mask, shift);
		uint64_t step = 0;
		for (; step <= mask; step++) {
			const unsigned long long old = atomicCAS(&tab[2 * pos], LP_EMPTY, key);
			if (old == LP_EMPTY) {
				tab[2 * pos + 1] = pay;
				break;
			}
			if (old == key) {
				dup = true;
				break;
			}
			pos = (pos + 1) & mask;
		}
		if (step > mask) full = true;
	}
	// one flag write per wave instead of one per duplicate
	if (__any(dup) && (threadIdx.x & (HJ_WAVE - 1)) == 0) atomicOr(&st->flags, HJ_DF_DUPKEY);
	if (__any(full) && (threadIdx.x & (HJ_WAVE - 1)) == 0) atomicOr(&st->flags, HJ_DF_TABLE_FULL);
}

// slot holding `key` (it must be present); used by the duplicate fix-up
template <bool MIX>
__device__ __forceinline__ uint64_t lp_find(const unsigned long long *tab, uint64_t mask, uint32_t shift, uint64_t key) {
	uint64_t pos = lp_home<MIX>(key, mask, shift);
	for (uint64_t step = 0; step <= mask; step++) {
		if (tab[2 * pos] == key) return pos;
		pos = (pos + 1) & mask;
	}
	return ~0ULL;
}

template <int WV, bool MIX>
__global__ __launch_bounds__(NPHJ_THREADS) void k_lp_minidx(const uint4 *__restrict__ R, uint64_t n,
                                                            const unsigned long long *__restrict__ tab, uint64_t mask,
                                                            uint32_t shift, unsigned long long *__restrict__ minidx) {
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
		const uint64_t key = hj_key_of(R[i * WV]);
		if (key == LP_EMPTY) continue;
		const uint64_t pos = lp_find<MIX>(tab, mask, shift, key);
		if (pos != ~0ULL) atomicMin(&minidx[pos], (unsigned long long)i);
	}
}

template <int WV, bool MIX>
__global__ __launch_bounds__(NPHJ_THREADS) void k_lp_setpay(const uint4 *__restrict__ R, uint64_t n,
                                                            unsigned long long *__restrict__ tab, uint64_t mask,
                                                            uint32_t shift, const unsigned long long *__restrict__ minidx) {
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
		const uint4 t = R[i * WV];
		const uint64_t key = hj_key_of(t);
		if (key == LP_EMPTY) continue;
		const uint64_t pos = lp_find<MIX>(tab, mask, shift, key);
		if (pos != ~0ULL && minidx[pos] == i) tab[2 * pos + 1] = hj_pay_of(t);
	}
}

template <int WV, bool EMIT, bool MIX>
__global__ __launch_bounds__(NPHJ_THREADS) void k_lp_probe(const uint4 *__restrict__ S, uint64_t n,
                                                           const ulonglong2 *__restrict__ tab, uint64_t mask,
                                                           uint32_t shift, const uint4 *__restrict__ R, hj_stats *st,
                                                           uint4 *pout,
                                                           unsigned long long pcap, unsigned pchunk) {
	__shared__ unsigned long long red[2 * (NPHJ_THREADS / HJ_WAVE)];
	__shared__ unsigned long long emit_ws[EMIT ? 2 * (NPHJ_THREADS / HJ_WAVE) : 2];
	auto acc = hj_make_sink<EMIT>(pout, pcap, &st->pair_next, pchunk, emit_ws);
	const uint64_t sent_cnt = st->sent_cnt;
	const uint64_t sent_pay = sent_cnt ? hj_pay_of(R[st->sent_min * WV]) : 0;
	const uint64_t tile = (uint64_t)blockDim.x * PROBE_U;
	for (uint64_t base = (uint64_t)blockIdx.x * tile; base < n; base += (uint64_t)gridDim.x * tile) {
		uint64_t key[PROBE_U], pay[PROBE_U], pos[PROBE_U];
		ulonglong2 e[PROBE_U];
		bool live[PROBE_U];
#pragma unroll
		for (int u = 0; u < PROBE_U; u++) {
			const uint64_t i = base + (uint64_t)u * blockDim.x + threadIdx.x;
			live[u] = i < n;
			uint4 t = live[u] ? hj_load_stream(&S[i * WV]) : make_uint4(0, 0, 0, 0);
			key[u] = hj_key_of(t);
			pay[u] = hj_pay_of(t);
		}
#pragma unroll
		for (int u = 0; u < PROBE_U; u++) {
			pos[u] = lp_home<MIX>(key[u], mask, shift);
			if (live[u] && key[u] != LP_EMPTY) e[u] = tab[pos[u]];
		}
#pragma unroll
		for (int u = 0; u < PROBE_U; u++) {
			if (!live[u]) continue;
			if (key[u] == LP_EMPTY) {
				if (sent_cnt) acc.match(sent_pay, pay[u]);
				continue;
			}
			for (uint64_t step = 0; step <= mask; step++) {
				if (e[u].x == key[u]) {
					acc.match(e[u].y, pay[u]);
					break;
				}
				if (e[u].x == LP_EMPTY) break;
				pos[u] = (pos[u] + 1) & mask;
				e[u] = tab[pos[u]];
			}
		}
	}
	acc.finish();
	hj_block_add2(acc.cnt, acc.sum, red, &st->matched, &st->checksum);
}

// ---- narrow LP table: one 8-byte slot {u32 key, u32 payload} per entry,
// claimed whole by one 64-bit CAS against the empty marker (all ones).  It
// takes every R whose keys are below 2^32 - 1 and whose payloads fit 32 bits
// (the build flags HJ_DF_WIDE on any other tuple and the host rebuilds the
// 16-byte table): half the bytes, so C2's 2^24 slots (128 MB) stay in the
// 256 MB Infinity Cache while S streams past, and a probe reads 8 bytes.
#define LPN_EMPTY 0xFFFFFFFFFFFFFFFFULL
__device__ __forceinline__ bool lpn_fits(uint64_t key, uint64_t pay) { return key < 0xFFFFFFFFULL && pay <= 0xFFFFFFFFULL; }

template <int WV, bool MIX>
__global__ __launch_bounds__(NPHJ_THREADS) void k_lpn_build(const uint4 *__restrict__ R, uint64_t n,
                                                            unsigned long long *__restrict__ tab, uint64_t mask,
                                                            uint32_t shift, hj_stats *st) {
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	bool dup = false, full = false, wide = false;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
		const uint4 t = R[i * WV];
		const uint64_t key = hj_key_of(t), pay = hj_pay_of(t);
		if (!lpn_fits(key, pay)) {
			wide = true;
			continue;
		}
		const unsigned long long v = key | pay << 32;
		uint64_t pos = lp_home<MIX>(key, mask, shift);
		uint64_t step = 0;
		for (; step <= mask; step++) {
			const unsigned long long old = atomicCAS(&tab[pos], LPN_EMPTY, v);
			if (old == LPN_EMPTY) break;
			if ((uint32_t)old == (uint32_t)key) {
				dup = true;
				break;
			}
			pos = (pos + 1) & mask;
		}
		if (step > mask) full = true;
	}
	const bool lane0 = (threadIdx.x & (HJ_WAVE - 1)) == 0;
	if (__any(dup) && lane0) atomicOr(&st->flags, HJ_DF_DUPKEY);
	if (__any(full) && lane0) atomicOr(&st->flags, HJ_DF_TABLE_FULL);
	if (__any(wide) && lane0) atomicOr(&st->flags, HJ_DF_WIDE);
}

// slot holding `key` (it must be present), narrow table
template <bool MIX>
__device__ __forceinline__ uint64_t lpn_find(const unsigned long long *tab, uint64_t mask, uint32_t shift, uint64_t key) {
	uint64_t pos = lp_home<MIX>(key, mask, shift);
	for (uint64_t step = 0; step <= mask; step++) {
		if ((uint32_t)tab[pos] == (uint32_t)key) return pos;
		pos = (pos + 1) & mask;
	}
	return ~0ULL;
}

template <int WV, bool MIX>
__global__ __launch_bounds__(NPHJ_THREADS) void k_lpn_minidx(const uint4 *__restrict__ R, uint64_t n,
                                                             const unsigned long long *__restrict__ tab, uint64_t mask,
                                                             uint32_t shift, unsigned long long *__restrict__ minidx) {
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
		const uint64_t pos = lpn_find<MIX>(tab, mask, shift, hj_key_of(R[i * WV]));
		if (pos != ~0ULL) atomicMin(&minidx[pos], (unsigned long long)i);
	}
}

template <int WV, bool MIX>
__global__ __launch_bounds__(NPHJ_THREADS) void k_lpn_setpay(const uint4 *__restrict__ R, uint64_t n,
                                                             unsigned long long *__restrict__ tab, uint64_t mask,
                                                             uint32_t shift, const unsigned long long *__restrict__ minidx) {
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
		const uint4 t = R[i * WV];
		const uint64_t key = hj_key_of(t);
		const uint64_t pos = lpn_find<MIX>(tab, mask, shift, key);
		if (pos != ~0ULL && minidx[pos] == i) tab[pos] = key | hj_pay_of(t) << 32;
	}
}

// Probe of the narrow table: every lane issues its PROBE_U first-slot loads
// together (clamped home slots, no lane condition), then walks on.  A probe
// key of 2^32 - 1 or more matches nothing (no such R key is in the table).
template <int WV, bool EMIT, bool MIX>
__global__ __launch_bounds__(NPHJ_THREADS) void k_lpn_probe(const uint4 *__restrict__ S, uint64_t n,
                                                            const unsigned long long *__restrict__ tab, uint64_t mask,
                                                            uint32_t shift, hj_stats *st, uint4 *pout,
                                                            unsigned long long pcap, unsigned pchunk) {
	__shared__ unsigned long long red[2 * (NPHJ_THREADS / HJ_WAVE)];
	__shared__ unsigned long long emit_ws[EMIT ? 2 * (NPHJ_THREADS / HJ_WAVE) : 2];
	auto acc = hj_make_sink<EMIT>(pout, pcap, &st->pair_next, pchunk, emit_ws);
	const uint64_t tile = (uint64_t)blockDim.x * PROBE_U;
	for (uint64_t base = (uint64_t)blockIdx.x * tile; base < n; base += (uint64_t)gridDim.x * tile) {
		uint64_t key[PROBE_U], pay[PROBE_U], pos[PROBE_U];
		unsigned long long e[PROBE_U];
#pragma unroll
		for (int u = 0; u < PROBE_U; u++) {
			const uint64_t i = min(base + (uint64_t)u * blockDim.x + threadIdx.x, n - 1);
			const uint4 t = hj_load_stream(&S[i * WV]);
			key[u] = hj_key_of(t);
			pay[u] = hj_pay_of(t);
		}
#pragma unroll
		for (int u = 0; u < PROBE_U; u++) {
			pos[u] = lp_home<MIX>(key[u], mask, shift);
			e[u] = tab[pos[u]];
		}
#pragma unroll
		for (int u = 0; u < PROBE_U; u++) {
			if (base + (uint64_t)u * blockDim.x + threadIdx.x >= n || key[u] >= 0xFFFFFFFFULL) continue;
			for (uint64_t step = 0; step <= mask; step++) {
				if ((uint32_t)e[u] == (uint32_t)key[u]) {
					acc.match(e[u] >> 32, pay[u]);
					break;
				}
				if (e[u] == LPN_EMPTY) break;
				pos[u] = (pos[u] + 1) & mask;
				e[u] = tab[pos[u]];
			}
		}
	}
	acc.finish();
	hj_block_add2(acc.cnt, acc.sum, red, &st->matched, &st->checksum);
}

// Build (and, if duplicates were seen, fix up) the LP table for R: the
// narrow table (8-byte slots) or the 16-byte one; `tab` has room for `slots`
// 16-byte slots (slots >= 2 when mix).  *wide_out: the narrow build met a
// tuple it cannot hold (the caller rebuilds the 16-byte table).
static int lp_build_table(hj_ctx_t *ctx, const hj_relation_t *R, unsigned long long *tab, uint64_t slots, bool mix,
                          bool narrow, bool *wide_out, hj_result_t *out) {
	const uint64_t mask = slots - 1;
	const uint32_t shift = 64 - hj_log2_ceil(slots);
	const uint32_t wv = R->tuple_size / 16;
	hj_stats *st = hj_stats_dev(ctx);
	const uint64_t n = R->tuple_num;
	const unsigned g = grid_for(ctx, n, NPHJ_THREADS);
	*wide_out = false;
	hj_kt_begin(ctx, HJ_K_LP_BUILD);
#define LP_B1(W, M)                                                                                               \
	if (narrow)                                                                                                   \
		hipLaunchKernelGGL((k_lpn_build<W, M>), dim3(g), dim3(NPHJ_THREADS), 0, ctx->stream, (const uint4 *)R->tuples, \
		                   n, tab, mask, shift, st);                                                              \
	else                                                                                                          \
		hipLaunchKernelGGL((k_lp_build<W, M>), dim3(g), dim3(NPHJ_THREADS), 0, ctx->stream, (const uint4 *)R->tuples,  \
		                   n, tab, mask, shift, st)
#define LP_B(W)          \
	if (mix) {           \
		LP_B1(W, true);  \
	} else {             \
		LP_B1(W, false); \
	}
	switch (wv) {
	case 1: LP_B(1); break;
	case 2: LP_B(2); break;
	case 4: LP_B(4); break;
	case 8: LP_B(8); break;
	case 16: LP_B(16); break;
	default: LP_B(32); break;
	}
#undef LP_B
#undef LP_B1
	hj_kt_end(ctx);
	HJ_HIP(ctx, hipGetLastError());
	// The duplicate flag is read back only to choose the fix-up path; PK
	// builds (unique keys) never take it.
	unsigned int flags = 0;
	HJ_HIP(ctx, hipMemcpyAsync(&flags, &st->flags, sizeof flags, hipMemcpyDeviceToHost, ctx->stream));
	HJ_HIP(ctx, hipStreamSynchronize(ctx->stream));
	if (narrow && (flags & HJ_DF_WIDE)) {
		*wide_out = true;
		return HJ_OK;
	}
	if (flags & HJ_DF_TABLE_FULL) return hj_fail(ctx, HJ_EINVAL, "nphj_lp table full: |R| exceeds slots");
	if (flags & HJ_DF_DUPKEY) {
		void *aux = nullptr;
		int rc = hj_scratch(ctx, ctx->aux, slots * 8, &aux);
		if (rc) return rc;
		hj_fill16(ctx, aux, (slots + 1) / 2, make_uint4(~0u, ~0u, ~0u, ~0u));
		hj_kt_begin(ctx, HJ_K_LP_FIXUP);
#define LP_F1(W, M)                                                                                               \
	if (narrow) {                                                                                                 \
		hipLaunchKernelGGL((k_lpn_minidx<W, M>), dim3(g), dim3(NPHJ_THREADS), 0, ctx->stream, (const uint4 *)R->tuples, \
		                   n, tab, mask, shift, (unsigned long long *)aux);                                       \
		hipLaunchKernelGGL((k_lpn_setpay<W, M>), dim3(g), dim3(NPHJ_THREADS), 0, ctx->stream, (const uint4 *)R->tuples, \
		                   n, tab, mask, shift, (const unsigned long long *)aux);                                 \
	} else {                                                                                                      \
		hipLaunchKernelGGL((k_lp_minidx<W, M>), dim3(g), dim3(NPHJ_THREADS), 0, ctx->stream, (const uint4 *)R->tuples,  \
		                   n, tab, mask, shift, (unsigned long long *)aux);                                       \
		hipLaunchKernelGGL((k_lp_setpay<W, M>), dim3(g), dim3(NPHJ_THREADS), 0, ctx->stream, (const uint4 *)R->tuples,  \
		                   n, tab, mask, shift, (const unsigned long long *)aux);                                 \
	}
#define LP_F(W)          \
	if (mix) {           \
		LP_F1(W, true);  \
	} else {             \
		LP_F1(W, false); \
	}
		switch (wv) {
		case 1: LP_F(1); break;
		case 2: LP_F(2); break;
		case 4: LP_F(4); break;
		case 8: LP_F(8); break;
		case 16: LP_F(16); break;
		default: LP_F(32); break;
		}
#undef LP_F
#undef LP_F1
		hj_kt_end(ctx);
		HJ_HIP(ctx, hipGetLastError());
		out->fallback |= HJ_FB_LP_DUPLICATES;
	}
	return HJ_OK;
}

static int lp_run(hj_ctx_t *ctx, const hj_relation_t *R, const hj_relation_t *S, uint64_t slots, bool mix,
                  hj_result_t *out) {
	void *tab_p = nullptr;
	int rc = hj_scratch(ctx, ctx->table, slots * 16, &tab_p);
	if (rc) return rc;
	unsigned long long *tab = (unsigned long long *)tab_p;
	hj_stats *st = hj_stats_dev(ctx);

	HJ_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
	hj_stats_reset(ctx);
	hj_kt_begin(ctx, HJ_K_FILL);
	hj_fill16(ctx, tab, (slots + 1) / 2, make_uint4(~0u, ~0u, ~0u, ~0u));  // narrow: 8-byte empty slots
	hj_kt_end(ctx);
	HJ_HIP(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
	bool narrow = true, wide = false;
	rc = lp_build_table(ctx, R, tab, slots, mix, true, &wide, out);
	if (rc) return rc;
	if (wide) {
		// a key of 2^32 - 1 or more, or a payload above 32 bits: the 16-byte
		// table (the counters start again; its fill is build time here)
		narrow = false;
		hj_stats_reset(ctx);
		hj_fill16(ctx, tab, slots, make_uint4(~0u, ~0u, 0u, 0u));
		rc = lp_build_table(ctx, R, tab, slots, mix, false, &wide, out);
		if (rc) return rc;
	}
	HJ_HIP(ctx, hipEventRecord(ctx->ev[2], ctx->stream));

	const uint32_t wv = S->tuple_size / 16;
	const uint64_t n = S->tuple_num;
	const unsigned g = grid_for(ctx, (n + PROBE_U - 1) / PROBE_U, NPHJ_THREADS, 8);
	hj_kt_begin(ctx, HJ_K_LP_PROBE);
#define LP_P1(W, E, M)                                                                                            \
	if (narrow)                                                                                                   \
		hipLaunchKernelGGL((k_lpn_probe<W, E, M>), dim3(g), dim3(NPHJ_THREADS), 0, ctx->stream,                     \
		                   (const uint4 *)S->tuples, n, (const unsigned long long *)tab, slots - 1,               \
		                   64 - hj_log2_ceil(slots), st, ctx->emit_pairs, (unsigned long long)ctx->emit_cap,      \
		                   ctx->emit_chunk);                                                                      \
	else                                                                                                          \
		hipLaunchKernelGGL((k_lp_probe<W, E, M>), dim3(g), dim3(NPHJ_THREADS), 0, ctx->stream,                      \
		                   (const uint4 *)S->tuples, n, (const ulonglong2 *)tab, slots - 1, 64 - hj_log2_ceil(slots), \
		                   (const uint4 *)R->tuples, st, ctx->emit_pairs, (unsigned long long)ctx->emit_cap,      \
		                   ctx->emit_chunk)
#define LP_P(W, E)          \
	if (mix) {              \
		LP_P1(W, E, true);  \
	} else {                \
		LP_P1(W, E, false); \
	}
	switch (wv) {
	case 1:
		if (ctx->emit_pairs) {
			LP_P(1, true);
		} else {
			LP_P(1, false);
		}
		break;
	case 2: LP_P(2, false); break;
	case 4: LP_P(4, false); break;
	case 8: LP_P(8, false); break;
	case 16: LP_P(16, false); break;
	default: LP_P(32, false); break;
	}
#undef LP_P
#undef LP_P1
	hj_kt_end(ctx);
	HJ_HIP(ctx, hipGetLastError());
	HJ_HIP(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
	hj_stats h;
	rc = hj_stats_read(ctx, &h);
	if (rc) return rc;
	out->matched_cnt = h.matched;
	out->check_sum = h.checksum;
	out->buildpart_time = hj_event_seconds(ctx, 1, 2);
	out->probejoin_time = hj_event_seconds(ctx, 2, 3);
	out->total_time = hj_event_seconds(ctx, 0, 3);
	// a table entry per tuple: 8 bytes in the narrow table, else 16
	out->narrow = narrow;
	out->algo_bytes = (uint64_t)(head_bytes(R->tuple_size) + (narrow ? 8 : 16)) * (R->tuple_num + S->tuple_num);
	return HJ_OK;
}

int hj_run_nphj_lp(hj_ctx_t *ctx, const hj_relation_t *R, const hj_relation_t *S, const hj_params_t *prm,
                   hj_result_t *out) {
	uint64_t slots = prm->lp_slots_log2 ? (1ULL << prm->lp_slots_log2) : hj_next_pow2(R->tuple_num ? R->tuple_num : 1);
	if (slots < R->tuple_num) return hj_fail(ctx, HJ_EINVAL, "lp_slots_log2 too small for |R|");
	return lp_run(ctx, R, S, slots, false, out);
}

// Exact first-match join for the lp fallbacks: nphj_lp's table and duplicate
// fix-up with the multiplicative home slot over next_pow_2(1.5 |R|) slots
// (at most 2/3 full), so its cost does not depend on the keys' low bits.
int hj_run_lp_exact(hj_ctx_t *ctx, const hj_relation_t *R, const hj_relation_t *S, hj_result_t *out) {
	const uint64_t n = R->tuple_num;
	return lp_run(ctx, R, S, std::max<uint64_t>(2, hj_next_pow2(n + n / 2)), true, out);
}

// ------------------------------------------------------------------ nphj_sc
// 64-byte bucket / overflow node: slots at u64 words 0..5 (key, payload)
// x3, u32 count at u32 word 12, u32 next (node index + 1) at u32 word 13.
#define SC_SLOTS 3

struct __align__(64) sc_node {
	unsigned long long kv[2 * SC_SLOTS];
	unsigned int count;
	unsigned int next;
	unsigned long long pad;
};

template <int WV>
__global__ __launch_bounds__(NPHJ_THREADS) void k_sc_build(const uint4 *__restrict__ R, uint64_t n,
                                                           sc_node *__restrict__ tab, uint64_t mask,
                                                           sc_node *__restrict__ pool, uint64_t pool_cap,
                                                           hj_stats *st) {
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	bool pool_full = false;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
		const uint4 t = R[i * WV];
		const uint64_t key = hj_key_of(t);
		sc_node *b = &tab[key & mask];
		const unsigned c = atomicAdd(&b->count, 1u);
		if (c < SC_SLOTS) {
			*(uint4 *)&b->kv[2 * c] = t;
			continue;
		}
		// overflow: append to the bucket's first node, or push a new node
		unsigned head = __hip_atomic_load(&b->next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		if (head) {
			const unsigned c2 = atomicAdd(&pool[head - 1].count, 1u);
			if (c2 < SC_SLOTS) {
				*(uint4 *)&pool[head - 1].kv[2 * c2] = t;
				continue;
			}
		}
		const unsigned long long nn = atomicAdd(&st->pool_next, 1ULL);
		if (nn >= pool_cap) {
			pool_full = true;
			continue;
		}
		sc_node *nd = &pool[nn];
		*(uint4 *)&nd->kv[0] = t;
		nd->count = 1u;
		__threadfence();
		// wait-free push: inserters only ever touch the head node, so the
		// link to the previous head can be stored after publication; the
		// probe kernel runs after this kernel has completed
		nd->next = atomicExch(&b->next, (unsigned)(nn + 1));
	}
	if (__any(pool_full) && (threadIdx.x & (HJ_WAVE - 1)) == 0) atomicOr(&st->flags, HJ_DF_POOL_FULL);
}

template <class Sink>
__device__ __forceinline__ void sc_match_node(const uint4 *nd, uint64_t key, uint64_t pay, Sink &sink,
                                              unsigned &next) {
	const uint4 w0 = nd[0], w1 = nd[1], w2 = nd[2], w3 = nd[3];
	const unsigned c = w3.x < SC_SLOTS ? w3.x : SC_SLOTS;
	next = w3.y;
	if (c > 0 && hj_key_of(w0) == key) sink.match(hj_pay_of(w0), pay);
	if (c > 1 && hj_key_of(w1) == key) sink.match(hj_pay_of(w1), pay);
	if (c > 2 && hj_key_of(w2) == key) sink.match(hj_pay_of(w2), pay);
}

template <int WV, bool EMIT>
__global__ __launch_bounds__(NPHJ_THREADS) void k_sc_probe(const uint4 *__restrict__ S, uint64_t n,
                                                           const sc_node *__restrict__ tab, uint64_t mask,
                                                           const sc_node *__restrict__ pool, uint64_t pool_cap,
                                                           hj_stats *st, uint4 *pout, unsigned long long pcap,
                                                           unsigned pchunk) {
	__shared__ unsigned long long red[2 * (NPHJ_THREADS / HJ_WAVE)];
	__shared__ unsigned long long emit_ws[EMIT ? 2 * (NPHJ_THREADS / HJ_WAVE) : 2];
	auto acc = hj_make_sink<EMIT>(pout, pcap, &st->pair_next, pchunk, emit_ws);
	const uint64_t tile = (uint64_t)blockDim.x * PROBE_U, gstep = (uint64_t)gridDim.x * tile;
	// Probe tuples of a tile and their bucket heads are loaded at clamped
	// indices, never under a lane condition, and the next tile's probe
	// tuples are issued after this tile's bucket loads and before its
	// matches, so a match waits only for its own bucket and the next tile's
	// tuples stay in flight (lanes past n are masked at the match).
	auto load_s = [&](uint4 (&t)[PROBE_U], uint64_t base) {
#pragma unroll
		for (int u = 0; u < PROBE_U; u++)
			t[u] = hj_load_stream(&S[min(base + (uint64_t)u * blockDim.x + threadIdx.x, n - 1) * WV]);
	};
	auto probe = [&](const uint4 (&t)[PROBE_U], uint64_t base, uint4 (&tn)[PROBE_U], uint64_t nbase) {
		uint4 w[PROBE_U][4];
#pragma unroll
		for (int u = 0; u < PROBE_U; u++) {
			const uint4 *nd = (const uint4 *)&tab[hj_key_of(t[u]) & mask];
#pragma unroll
			for (int q = 0; q < 4; q++) w[u][q] = nd[q];
		}
		__builtin_amdgcn_sched_barrier(0);  // keep the bucket loads ahead of the next tile's
		load_s(tn, nbase);
		__builtin_amdgcn_sched_barrier(0);
#pragma unroll
		for (int u = 0; u < PROBE_U; u++) {
			if (base + (uint64_t)u * blockDim.x + threadIdx.x >= n) continue;
			const uint64_t key = hj_key_of(t[u]), pay = hj_pay_of(t[u]);
			const unsigned c = w[u][3].x < SC_SLOTS ? w[u][3].x : SC_SLOTS;
			if (c > 0 && hj_key_of(w[u][0]) == key) acc.match(hj_pay_of(w[u][0]), pay);
			if (c > 1 && hj_key_of(w[u][1]) == key) acc.match(hj_pay_of(w[u][1]), pay);
			if (c > 2 && hj_key_of(w[u][2]) == key) acc.match(hj_pay_of(w[u][2]), pay);
			unsigned next = w[u][3].y;
			uint64_t hops = 0;
			while (next && hops++ < pool_cap) {
				unsigned nx;
				sc_match_node((const uint4 *)&pool[next - 1], key, pay, acc, nx);
				next = nx;
			}
		}
	};
	const uint64_t b0 = (uint64_t)blockIdx.x * tile;
	if (b0 < n) {
		uint4 ta[PROBE_U], tb[PROBE_U];
		load_s(ta, b0);
		for (uint64_t base = b0; base < n; base += 2 * gstep) {
			probe(ta, base, tb, base + gstep);
			probe(tb, base + gstep, ta, base + 2 * gstep);  // past n: every lane masked
		}
	}
	acc.finish();
	hj_block_add2(acc.cnt, acc.sum, red, &st->matched, &st->checksum);
}

int hj_run_nphj_sc(hj_ctx_t *ctx, const hj_relation_t *R, const hj_relation_t *S, const hj_params_t *prm,
                   hj_result_t *out) {
	(void)prm;
	// |R|/2 buckets as src/algo/nphj.c:746 (BUCKET_CAPACITY 2), rounded to a
	// power of two so that key & mask is a valid identity hash.
	uint64_t nb = hj_next_pow2(R->tuple_num ? R->tuple_num : 1) / 2;
	if (nb == 0) nb = 1;
	uint64_t pool_cap = R->tuple_num / 4 + 1024;
	// a no-allocation join takes the pool that never fills (one node per
	// build tuple) from its fixed region at once, so it never rebuilds
	if (ctx->fixed) pool_cap = R->tuple_num + 1024;
	hj_stats *st = hj_stats_dev(ctx);
	const uint32_t wv = R->tuple_size / 16;
	void *tab_p = nullptr, *pool_p = nullptr;
	int rc = hj_scratch(ctx, ctx->table, nb * sizeof(sc_node), &tab_p);
	if (rc) return rc;
	for (int attempt = 0; attempt < 2; attempt++) {
		if (attempt == 0 || !ctx->fixed) rc = hj_scratch(ctx, ctx->pool, pool_cap * sizeof(sc_node), &pool_p);
		if (rc) return rc;
		sc_node *tab = (sc_node *)tab_p, *pool = (sc_node *)pool_p;
		HJ_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
		hj_stats_reset(ctx);
		hj_kt_begin(ctx, HJ_K_FILL);
		hj_fill16(ctx, tab, nb * 4, make_uint4(0, 0, 0, 0));
		hj_kt_end(ctx);
		HJ_HIP(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
		const unsigned g = grid_for(ctx, R->tuple_num, NPHJ_THREADS);
		hj_kt_begin(ctx, HJ_K_SC_BUILD);
#define SC_B(W)                                                                                                   \
	hipLaunchKernelGGL(k_sc_build<W>, dim3(g), dim3(NPHJ_THREADS), 0, ctx->stream, (const uint4 *)R->tuples, \
	                   R->tuple_num, tab, nb - 1, pool, pool_cap, st)
		switch (wv) {
		case 1: SC_B(1); break;
		case 2: SC_B(2); break;
		case 4: SC_B(4); break;
		case 8: SC_B(8); break;
		case 16: SC_B(16); break;
		default: SC_B(32); break;
		}
#undef SC_B
		hj_kt_end(ctx);
		HJ_HIP(ctx, hipGetLastError());
		HJ_HIP(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
		const uint64_t n = S->tuple_num;
		const unsigned gp = grid_for(ctx, (n + PROBE_U - 1) / PROBE_U, NPHJ_THREADS, 8);
		hj_kt_begin(ctx, HJ_K_SC_PROBE);
#define SC_P(W, E)                                                                                                   \
	hipLaunchKernelGGL((k_sc_probe<W, E>), dim3(gp), dim3(NPHJ_THREADS), 0, ctx->stream, (const uint4 *)S->tuples, n, \
	                   (const sc_node *)tab, nb - 1, (const sc_node *)pool, pool_cap, st, ctx->emit_pairs,          \
	                   (unsigned long long)ctx->emit_cap, ctx->emit_chunk)
		switch (wv) {
		case 1:
			if (ctx->emit_pairs)
				SC_P(1, true);
			else
				SC_P(1, false);
			break;
		case 2: SC_P(2, false); break;
		case 4: SC_P(4, false); break;
		case 8: SC_P(8, false); break;
		case 16: SC_P(16, false); break;
		default: SC_P(32, false); break;
		}
#undef SC_P
		hj_kt_end(ctx);
		HJ_HIP(ctx, hipGetLastError());
		HJ_HIP(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
		hj_stats h;
		rc = hj_stats_read(ctx, &h);
		if (rc) return rc;
		if (h.flags & HJ_DF_POOL_FULL) {
			// every node holds at least one tuple, so |R| nodes always suffice
			pool_cap = R->tuple_num + 1024;
			out->fallback |= HJ_FB_SC_REBUILD;
			ctx->n_kt = 0;
			continue;
		}
		out->matched_cnt = h.matched;
		out->check_sum = h.checksum;
		out->buildpart_time = hj_event_seconds(ctx, 1, 2);
		out->probejoin_time = hj_event_seconds(ctx, 2, 3);
		out->total_time = hj_event_seconds(ctx, 0, 3);
		out->algo_bytes = (uint64_t)(head_bytes(R->tuple_size) + 16) * (R->tuple_num + S->tuple_num);
		return HJ_OK;
	}
	return hj_fail(ctx, HJ_ENOMEM, "nphj_sc overflow pool exhausted after rebuild");
}
