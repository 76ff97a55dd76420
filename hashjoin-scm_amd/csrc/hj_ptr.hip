// hj_ptr.hip -- late materialisation: the ptr_* joins.
//
// The reference's ptr_ algorithms (src/algo/ptr_phj.c, ptr_nphj.c,
// ptr_phj_bw_reg.c) join only the 16-byte {key, row_id} head of every tuple
// (keyrid_t, src/inc/store.h:81-84), write one row_id_pair_t {r.row_id,
// s.row_id} per match (store.h:99-102; e.g. ptr_phj.c:753-763), and then a
// retrieve phase dereferences each pair back into the original relations,
// adding R[r_row_id - 1].row_id + S[s_row_id - 1].row_id
// (ptr_retrieve, src/inc/utils.c:712-723).  The printed check_sum is that
// retrieve sum; matched_cnt is the pair count.
//
// Here:
//   1. tuples wider than 16 bytes have their heads copied into dense
//      16-byte relations (one streaming pass; the reference does this inside
//      its first partitioning pass, ptr_phj.c:218);
//   2. the ordinary join engine runs on those heads with the context's pair
//      buffer armed, so its probe kernels append every (build row_id, probe
//      row_id) pair (hj_pair_sink);
//   3. k_retrieve reads the pairs and gathers both row ids.
// A row id outside 1..|R| (or 1..|S|) is reported as an error: the reference
// reads out of bounds there.  Pairs are written only for matches; the
// reference's ptr_lp_join with NONTEMP_STORE_MATER also writes a zero pair
// per probe miss (ptr_phj.c:871-874), which its retrieve would dereference
// at row -1 -- pkfk inputs, the only ones it runs, have no misses.
// Probe kernels claim pair slots in per-wave chunks (one global atomic per
// chunk) and mark the unused tail of each chunk as a hole.  The buffer is
// sized for |S| pairs (pk-fk joins match each probe tuple once) plus a
// hole allowance of one chunk per resident wave; a join that emits more
// runs again with the exact count plus that allowance.

#include "hj_device.h"
#include "hj_internal.h"

#include <algorithm>

#define PT_THREADS 256
#define KR_U 4
#define RT_U 8

// heads of n tuples of `wv` 16-byte words into a dense 16-byte relation
__global__ __launch_bounds__(PT_THREADS) void k_keyrid(const uint4 *__restrict__ in, uint64_t n, uint32_t wv,
                                                       uint4 *__restrict__ out) {
	const uint64_t tile = (uint64_t)PT_THREADS * KR_U;
	for (uint64_t base = (uint64_t)blockIdx.x * tile; base < n; base += (uint64_t)gridDim.x * tile) {
		uint4 v[KR_U];
		// unconditional loads (clamped to the last tuple): a load under a
		// per-lane condition waits for its data before the next one issues
#pragma unroll
		for (int u = 0; u < KR_U; u++) v[u] = hj_load_stream(in + min(base + (uint64_t)u * PT_THREADS + threadIdx.x, n - 1) * wv);
#pragma unroll
		for (int u = 0; u < KR_U; u++) {
			const uint64_t i = base + (uint64_t)u * PT_THREADS + threadIdx.x;
			if (i < n) hj_store_stream(out + i, v[u]);
		}
	}
}

// Sum of R[r - 1].row_id + S[s - 1].row_id over n pair slots {r, s},
// skipping holes (both words HJ_PAIR_HOLE).  Row ids are 1-based; a pair
// with r outside 1..nr or s outside 1..ns adds nothing.  Adds the sum to
// st->retrieved and the number of in-range pairs to st->retrieved_ok.  R and
// S are seen as u64 words, w8 per tuple, row_id at word 1.  Each lane keeps
// 2 * RT_U gathers in flight.
__global__ __launch_bounds__(PT_THREADS) void k_retrieve(const uint4 *__restrict__ pairs, uint64_t n,
                                                         const uint64_t *__restrict__ R, uint64_t nr,
                                                         const uint64_t *__restrict__ S, uint64_t ns, uint32_t w8,
                                                         hj_stats *st) {
	__shared__ unsigned long long red[2 * (PT_THREADS / HJ_WAVE)];
	unsigned long long sum = 0, good = 0;
	const uint64_t tile = (uint64_t)PT_THREADS * RT_U;
	const HJ_GLOBAL uint64_t *gR = hj_g(R), *gS = hj_g(S);
	for (uint64_t base = (uint64_t)blockIdx.x * tile; base < n; base += (uint64_t)gridDim.x * tile) {
		uint64_t ri[RT_U], si[RT_U];
		bool ok[RT_U];
#pragma unroll
		for (int u = 0; u < RT_U; u++) {
			const uint64_t i = base + (uint64_t)u * PT_THREADS + threadIdx.x;
			const bool live = i < n;
			const uint4 p = hj_load_stream(pairs + min(i, n - 1));  // unconditional, clamped
			const uint64_t r = hj_key_of(p) - 1, s = hj_pay_of(p) - 1;  // 0 wraps to out of range
			ok[u] = live && r < nr && s < ns;  // a hole wraps to ~0 - 1: out of range
			good += ok[u];
			// an out-of-range pair reads row 0's row id instead (n > 0 means
			// both relations have a row) and adds nothing
			ri[u] = (ok[u] ? r : 0) * w8 + 1;
			si[u] = (ok[u] ? s : 0) * w8 + 1;
		}
		uint64_t a[RT_U], b[RT_U];
#pragma unroll
		for (int u = 0; u < RT_U; u++) {
			a[u] = gR[ri[u]];
			b[u] = gS[si[u]];
		}
#pragma unroll
		for (int u = 0; u < RT_U; u++) sum += ok[u] ? a[u] + b[u] : 0;
	}
	hj_block_add2(sum, good, red, &st->retrieved, &st->retrieved_ok);
}

static unsigned pt_grid(hj_ctx_t *ctx, uint64_t n, uint64_t per_block) {
	const uint64_t want = (n + per_block - 1) / per_block;
	return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(want, (uint64_t)ctx->num_cu * 8));
}

void hj_launch_keyrid(hj_ctx_t *ctx, const void *in, uint64_t n, uint32_t wv, void *out) {
	if (n)
		hipLaunchKernelGGL(k_keyrid, dim3(pt_grid(ctx, n, PT_THREADS * KR_U)), dim3(PT_THREADS), 0, ctx->stream,
		                   (const uint4 *)in, n, wv, (uint4 *)out);
}

// Disarms the pair buffer whatever way the inner join returns.
struct emit_guard {
	hj_ctx_t *ctx;
	~emit_guard() {
		ctx->emit_pairs = nullptr;
		ctx->emit_cap = 0;
	}
};

int hj_run_ptr(hj_ctx_t *ctx, const hj_relation_t *R, const hj_relation_t *S, const hj_params_t *prm,
               hj_algo_fn inner, hj_result_t *out) {
	const uint32_t W = R->tuple_size, wv = W / 16;
	const uint64_t nr = R->tuple_num, ns = S->tuple_num;
	const int e0 = hj_ctx::EV_PTR;
	hj_stats *st = (hj_stats *)ctx->stats.p;
	int rc;

	// 1. heads of wide tuples.  A radix join that takes two passes reads them
	// in place in its first pass, and with one pass copies them itself;
	// the NPHJ probes that emit pairs read dense 16-byte heads.
	hj_relation_t kr = *R, ks = *S;
	const bool nphj = inner == hj_run_nphj_sc || inner == hj_run_nphj_lp;
	HJ_HIP(ctx, hipEventRecord(ctx->ev[e0], ctx->stream));
	if (wv > 1 && nphj) {
		if ((rc = hj_reserve(ctx, ctx->keyrid[0], nr * 16)) || (rc = hj_reserve(ctx, ctx->keyrid[1], ns * 16)))
			return rc;
		hj_kt_begin(ctx, HJ_K_KEYRID);
		hj_launch_keyrid(ctx, R->tuples, nr, wv, ctx->keyrid[0].p);
		hj_launch_keyrid(ctx, S->tuples, ns, wv, ctx->keyrid[1].p);
		hj_kt_end(ctx);
		HJ_HIP(ctx, hipGetLastError());
		kr.tuples = ctx->keyrid[0].p;
		ks.tuples = ctx->keyrid[1].p;
		kr.tuple_size = ks.tuple_size = 16;
	}
	HJ_HIP(ctx, hipEventRecord(ctx->ev[e0 + 1], ctx->stream));

	// 2. join emitting row-id pairs.  Every emitting kernel runs at most 32
	// waves per CU, each leaving less than one chunk of holes; the chunk is
	// sized so that the holes take about a quarter of |S| (at most 8192
	// slots: ~M / 8192 atomics for M pairs).
	const uint64_t waves = (uint64_t)ctx->num_cu * 32;
	uint64_t chunk = 256;
	while (chunk < 8192 && waves * chunk * 2 * 4 <= ns) chunk *= 2;
	uint64_t cap = std::max<uint64_t>(ns + waves * chunk, ctx->pairs.cap / 16);
	const int kt_mark = ctx->n_kt;
	hj_result_t r{};
	hj_stats h{};
	for (int attempt = 0;; attempt++) {
		if ((rc = hj_reserve(ctx, ctx->pairs, cap * 16))) return rc;
		cap = ctx->pairs.cap / 16;
		{
			emit_guard g{ctx};
			ctx->emit_pairs = (uint4 *)ctx->pairs.p;
			ctx->emit_cap = cap;
			ctx->emit_chunk = (unsigned)chunk;
			r = hj_result_t{};
			if ((rc = inner(ctx, &kr, &ks, prm, &r))) return rc;
		}
		if ((rc = hj_stats_read(ctx, &h))) return rc;
		if (h.pair_next <= cap) break;
		if (attempt >= 2) return hj_fail(ctx, HJ_EHIP, "pair buffer overflow after growing to the match count");
		cap = std::max<uint64_t>(h.matched, r.matched_cnt) + waves * chunk;
		ctx->n_kt = kt_mark;
	}
	const double t_extract = hj_event_seconds(ctx, e0, e0 + 1);

	// 3. retrieve over the claimed slots (the inner join's stats reset
	// cleared retrieved / retrieved_ok)
	const uint64_t m = h.pair_next;
	HJ_HIP(ctx, hipEventRecord(ctx->ev[e0], ctx->stream));
	hj_kt_begin(ctx, HJ_K_RETRIEVE);
	if (m)
		hipLaunchKernelGGL(k_retrieve, dim3(pt_grid(ctx, m, PT_THREADS * RT_U)), dim3(PT_THREADS), 0, ctx->stream,
		                   (const uint4 *)ctx->pairs.p, m, (const uint64_t *)R->tuples, nr,
		                   (const uint64_t *)S->tuples, ns, W / 8, st);
	hj_kt_end(ctx);
	HJ_HIP(ctx, hipGetLastError());
	HJ_HIP(ctx, hipEventRecord(ctx->ev[e0 + 2], ctx->stream));
	if ((rc = hj_stats_read(ctx, &h))) return rc;
	if (h.retrieved_ok != r.matched_cnt)
		return hj_fail(ctx, HJ_EINVAL,
		               "late materialisation: %llu of %llu row-id pairs point outside R or S (row_id must be "
		               "the 1-based position of a tuple in its relation)",
		               (unsigned long long)(r.matched_cnt - h.retrieved_ok), (unsigned long long)r.matched_cnt);
	const double t_retrieve = hj_event_seconds(ctx, e0, e0 + 2);

	*out = r;
	out->check_sum = h.retrieved;
	out->buildpart_time = t_extract + r.buildpart_time;
	out->probejoin_time = r.probejoin_time;
	out->total_time = t_extract + r.total_time + t_retrieve;
	// heads read (one burst of up to 64 bytes) and written once, pairs written
	// and read once, two 8-byte row ids gathered per pair
	// (a radix inner join counts its own head reads)
	out->algo_bytes = r.algo_bytes + (wv > 1 && nphj ? (uint64_t)(std::min<uint32_t>(W, 64) + 16) * (nr + ns) : 0) +
	                  48 * r.matched_cnt;
	return HJ_OK;
}
