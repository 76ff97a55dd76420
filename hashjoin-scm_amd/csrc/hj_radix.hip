// hj_radix.hip -- radix-partitioned hash join (phj_radix_bc / phj_radix_lp).
//
// Reference: src/algo/phj.c -- join_init_run (:1875-2512), partitioning
// (:147-359 one pass, :366-491 second pass), bc_join (:496-571), lp_join
// (:574-680), task queue src/inc/task_pthread.h.
//
// Partitioning (one "multi-split" per pass).  The input of a pass is a list
// of segments (pass 1: R and S; pass 2: every pass-1 partition), each cut
// into chunks of CH tuples; one workgroup owns one chunk, the way one
// reference thread owns one slice (phj.c:2046-2053).
//   k_hist     per-chunk histogram of the radix digit in LDS, written to a
//              flat matrix laid out [segment][digit][chunk];
//   k_scan     one exclusive scan over the flat matrix gives every
//              (segment, digit, chunk) its absolute output position -- the
//              reference's cross-thread prefix (phj.c:299-311) for all
//              segments at once;
//   k_scatter  re-reads the chunk tile by tile; each tile is counting-sorted
//              by digit in LDS and written out as contiguous runs per digit,
//              so a wavefront stores whole lines instead of one tuple per
//              partition.  This LDS staging replaces the reference's 64-byte
//              software write-combining buffers + AVX-512 non-temporal
//              flushes (phj.c:218-254, :314-358).
// Digit bits follow the reference: pass 1 uses key bits [0, B/P), pass 2
// bits [B/P, B) (params.h:34-35); the in-partition hash uses key >> B.
//
// Join.  A planner turns every final partition pair (R_q, S_q) into tiles of
// at most s_tile probe tuples; a persistent kernel pulls (partition, tile)
// items from a device counter -- the GPU form of the reference's LIFO task
// queue (task_pthread.h:110-133) -- builds the R_q table in LDS and probes
// the tile.  bc: bucket chains (heads + next, 1-based, head insertion, as
// bc_join).  lp: linear probing over next_pow_2(|R_q|) slots, first match
// only (as lp_join).  Count and checksum are wave-reduced and added with one
// 64-bit atomic per workgroup.

#include "hj_device.h"
#include "hj_internal.h"
#include "hj_tables.h"

#include <algorithm>
#include <string.h>
#include <vector>


struct hj_seg {
	const uint4 *in;     // first input tuple
	uint64_t len;        // tuples
	uint64_t out_start;  // absolute output position of the segment
	uint64_t hist_base;  // flat index of (digit 0, chunk 0)
	uint32_t chunk_base; // first global chunk id
	uint32_t nchunks;
	uint32_t bits;       // radix digit width of this segment
	uint32_t pb_base;    // first entry of its partition bounds (batches)
	uint32_t in_wv;      // 16-byte words per input tuple when pass A reads only
	                     // the heads of wider tuples (0: the kernel's own width)
	uint32_t pad;
};

struct hj_pair {
	uint64_t r0, r1, s0, s1;  // tuple ranges in the partitioned buffer
};

#define SCAN_NT 256
#define SCAN_ITEMS 16
#define SCAN_CH (SCAN_NT * SCAN_ITEMS)

// ------------------------------------------------------------------ scans
// In-place exclusive scan of a[0..L) (u64); a[L] receives the total.
template <typename T>
__global__ __launch_bounds__(SCAN_NT) void k_scan_reduce(const T *__restrict__ a, uint64_t L,
                                                         uint64_t *__restrict__ part) {
	__shared__ uint64_t red[SCAN_NT / HJ_WAVE + 1];
	const uint64_t base = (uint64_t)blockIdx.x * SCAN_CH;
	// unconditional loads (clamped, masked after): a load under a per-lane
	// condition waits for its data before the next one issues.  a holds L + 1
	// entries, so the clamp stays inside it when L is 0.
	T v[SCAN_ITEMS];
#pragma unroll
	for (int k = 0; k < SCAN_ITEMS; k++) v[k] = a[min(base + (uint64_t)k * SCAN_NT + threadIdx.x, L ? L - 1 : 0)];
	uint64_t s = 0;
#pragma unroll
	for (int k = 0; k < SCAN_ITEMS; k++) s += base + (uint64_t)k * SCAN_NT + threadIdx.x < L ? v[k] : 0;
	uint64_t tot;
	hj_block_exscan<uint64_t>(s, red, &tot);
	if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

template <typename T>
__global__ __launch_bounds__(1024) void k_scan_top(uint64_t *__restrict__ part, uint32_t nparts,
                                                   T *__restrict__ total_out) {
	__shared__ uint64_t red[1024 / HJ_WAVE + 1];
	// each thread owns a contiguous run of partials
	const uint32_t per = (nparts + blockDim.x - 1) / blockDim.x;
	const uint32_t b0 = threadIdx.x * per;
	uint64_t s = 0;
	for (uint32_t k = 0; k < per; k++)
		if (b0 + k < nparts) s += part[b0 + k];
	uint64_t tot;
	uint64_t run = hj_block_exscan<uint64_t>(s, red, &tot);
	for (uint32_t k = 0; k < per; k++)
		if (b0 + k < nparts) {
			uint64_t t = part[b0 + k];
			part[b0 + k] = run;
			run += t;
		}
	if (threadIdx.x == 0) *total_out = (T)tot;
}

// The block's SCAN_CH elements move between HBM and LDS coalesced (element
// k * SCAN_NT + t by thread t) and each thread scans its SCAN_ITEMS
// consecutive elements in LDS.
template <typename T>
__global__ __launch_bounds__(SCAN_NT) void k_scan_down(T *__restrict__ a, uint64_t L,
                                                       const uint64_t *__restrict__ part) {
	__shared__ uint64_t red[SCAN_NT / HJ_WAVE + 1];
	// row pitch SCAN_ITEMS + 1 words: a thread's consecutive elements sit in
	// consecutive banks and the threads of a wave start in different banks
	__shared__ uint64_t buf[SCAN_NT * (SCAN_ITEMS + 1)];
	const uint64_t base = (uint64_t)blockIdx.x * SCAN_CH;
	auto at = [](uint32_t e) { return (e / SCAN_ITEMS) * (SCAN_ITEMS + 1) + e % SCAN_ITEMS; };
	T g[SCAN_ITEMS];
#pragma unroll
	for (int k = 0; k < SCAN_ITEMS; k++) g[k] = a[min(base + (uint64_t)k * SCAN_NT + threadIdx.x, L ? L - 1 : 0)];
#pragma unroll
	for (int k = 0; k < SCAN_ITEMS; k++) {
		const uint32_t e = (uint32_t)k * SCAN_NT + threadIdx.x;
		buf[at(e)] = base + e < L ? g[k] : 0;
	}
	__syncthreads();
	uint64_t v[SCAN_ITEMS];
	uint64_t s = 0;
#pragma unroll
	for (int k = 0; k < SCAN_ITEMS; k++) {
		v[k] = buf[threadIdx.x * (SCAN_ITEMS + 1) + k];
		s += v[k];
	}
	uint64_t tot;
	uint64_t run = hj_block_exscan<uint64_t>(s, red, &tot) + part[blockIdx.x];
#pragma unroll
	for (int k = 0; k < SCAN_ITEMS; k++) {
		buf[threadIdx.x * (SCAN_ITEMS + 1) + k] = run;
		run += v[k];
	}
	__syncthreads();
#pragma unroll
	for (int k = 0; k < SCAN_ITEMS; k++) {
		const uint32_t e = (uint32_t)k * SCAN_NT + threadIdx.x;
		if (base + e < L) a[base + e] = (T)buf[at(e)];
	}
}

template <typename T>
__global__ void k_copy_total(const uint64_t *__restrict__ src, T *__restrict__ dst) { *dst = (T)*src; }

// Exclusive scan of a[0..L) in place, a[L] = total.  `tmp` holds the
// ceil(L / SCAN_CH) partials; beyond 65536 partials they are scanned the
// same way one level up, which needs ceil(nb / SCAN_CH) + 2 more entries
// (callers reserve partials + 2 + 512 entries: L up to 2^33).
template <typename T>
static int scan_t(hj_ctx_t *ctx, T *a, uint64_t L, uint64_t *tmp) {
	uint64_t nb = (L + SCAN_CH - 1) / SCAN_CH;
	if (nb == 0) nb = 1;
	if (nb > ((uint64_t)1 << 31) - 1) return hj_fail(ctx, HJ_EINVAL, "scan too long (%llu)", (unsigned long long)L);
	hipLaunchKernelGGL(k_scan_reduce<T>, dim3((unsigned)nb), dim3(SCAN_NT), 0, ctx->stream, a, L, tmp);
	if (nb <= 1024 * 64) {
		hipLaunchKernelGGL(k_scan_top<T>, dim3(1), dim3(1024), 0, ctx->stream, tmp, (uint32_t)nb, a + L);
	} else {
		const int rc = scan_t<uint64_t>(ctx, tmp, nb, tmp + nb + 1);
		if (rc) return rc;
		hipLaunchKernelGGL(k_copy_total<T>, dim3(1), dim3(1), 0, ctx->stream, tmp + nb, a + L);
	}
	hipLaunchKernelGGL(k_scan_down<T>, dim3((unsigned)nb), dim3(SCAN_NT), 0, ctx->stream, a, L, tmp);
	return HJ_OK;
}
static int scan_u64(hj_ctx_t *ctx, uint64_t *a, uint64_t L, uint64_t *tmp) { return scan_t<uint64_t>(ctx, a, L, tmp); }
// the two-pass count matrix: positions below 2^32 (relations of at most 2^31
// tuples each)
static int scan_u32(hj_ctx_t *ctx, uint32_t *a, uint64_t L, uint64_t *tmp) { return scan_t<uint32_t>(ctx, a, L, tmp); }

// ------------------------------------------------------------ segments
// Pass 1: R and S as two segments; chunk map filled for both.
__global__ void k_seg_pair(const uint4 *R, uint64_t nr, const uint4 *S, uint64_t ns, uint64_t ch, uint32_t F,
                           hj_seg *segs, uint32_t *chunk_seg, hj_stats *st, uint32_t in_wv) {
	const uint32_t cr = (uint32_t)((nr + ch - 1) / ch), cs = (uint32_t)((ns + ch - 1) / ch);
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		const uint32_t bits = 31 - __clz(F);
		segs[0] = hj_seg{R, nr, 0, 0, 0, cr, bits, 0, in_wv, 0};
		segs[1] = hj_seg{S, ns, nr, (uint64_t)cr * F, cr, cs, bits, F + 1, in_wv, 0};
		st->total_chunks = cr + cs;
	}
	for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < cr + cs; c += gridDim.x * blockDim.x)
		chunk_seg[c] = c < cr ? 0u : 1u;
}

// Partition bounds: pb[s*F + p] = first output position of digit p of
// segment s; pb[nseg*F] = end of the last segment.
template <typename T>
__global__ void k_bounds(const hj_seg *__restrict__ segs, uint32_t nseg, uint32_t F, const T *__restrict__ scan,
                         uint64_t *__restrict__ pb) {
	const uint64_t total = (uint64_t)nseg * F;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= total;
	     i += (uint64_t)gridDim.x * blockDim.x) {
		if (i == total) {
			const hj_seg &l = segs[nseg - 1];
			pb[i] = l.out_start + l.len;
			continue;
		}
		const uint32_t s = (uint32_t)(i / F), p = (uint32_t)(i % F);
		const hj_seg &g = segs[s];
		pb[i] = g.nchunks ? scan[g.hist_base + (uint64_t)p * g.nchunks] : g.out_start;
	}
}

// ------------------------------------------------------------ histogram
// DETECT (16-byte tuples whose pass may write 8-byte ones): also flag
// HJ_DF_WIDE when a key or row id of the chunk needs more than 32 bits; the
// loads are then whole tuples, which read the same lines.
template <int WV, bool DETECT>
__global__ __launch_bounds__(512) void k_hist(const hj_seg *__restrict__ segs, const uint32_t *__restrict__ chunk_seg,
                                              hj_stats *__restrict__ st, uint64_t ch, uint32_t shift,
                                              uint32_t bits, uint64_t *__restrict__ hist) {
	extern __shared__ uint32_t cnt[];
	const uint32_t c = blockIdx.x;
	if (c >= st->total_chunks) return;
	const hj_seg g = segs[chunk_seg[c]];
	const uint32_t F = 1u << min(bits, g.bits), fmask = F - 1;  // `bits` = launch maximum (LDS size)
	const uint32_t cl = c - g.chunk_base;
	const uint64_t b = (uint64_t)cl * ch, e = min(g.len, b + ch);
	for (uint32_t p = threadIdx.x; p < F; p += blockDim.x) cnt[p] = 0;
	__syncthreads();
	const uint64_t *keys = (const uint64_t *)g.in;
	constexpr int U = 8;
	bool wide = false;
	for (uint64_t i0 = b + threadIdx.x; i0 < e; i0 += (uint64_t)U * blockDim.x) {
		uint64_t k[U];
		if constexpr (DETECT) {
			uint4 t[U];
#pragma unroll
			for (int u = 0; u < U; u++) t[u] = hj_load_stream(g.in + min(i0 + (uint64_t)u * blockDim.x, e - 1));
#pragma unroll
			for (int u = 0; u < U; u++) {
				k[u] = hj_key_of(t[u]);
				wide |= (t[u].y | t[u].w) != 0;  // past the end: the last tuple again
			}
		} else {
#pragma unroll
			for (int u = 0; u < U; u++) {
				// clamped, never under a lane condition (masked at the count)
				const uint64_t i = min(i0 + (uint64_t)u * blockDim.x, e - 1);
				k[u] = __builtin_nontemporal_load(hj_g(&keys[i * 2 * WV]));
			}
		}
#pragma unroll
		for (int u = 0; u < U; u++) {
			const uint64_t i = i0 + (uint64_t)u * blockDim.x;
			if (i < e) atomicAdd(&cnt[(uint32_t)(k[u] >> shift) & fmask], 1u);
		}
	}
	if (DETECT && __syncthreads_or(wide) && threadIdx.x == 0) atomicOr(&st->flags, HJ_DF_WIDE);
	__syncthreads();
	for (uint32_t p = threadIdx.x; p < F; p += blockDim.x) hist[g.hist_base + (uint64_t)p * g.nchunks + cl] = cnt[p];
}

// ------------------------------------------------------------ scatter
// Each tile is counting-sorted by digit in LDS and written out as runs.  A
// run's tail that does not fill a `lt`-tuple line (lt * W = 64 or 128 bytes)
// is held back in LDS (`carry`) and written with the next tile's run of the
// same digit, so every store of a digit's stream after its first covers whole
// lines; the carries are flushed when the chunk ends.  This is the GPU form
// of software write-combining buffers: partial-line writes cost as much HBM
// time as full ones.  lt = 1 disables combining.
// LDS: tile[XT*WV] uint4 | carry[FL*(lt-1)*WV] uint4 | cursor[FL] u64 |
//      tcnt, tbase, cf [FL] u32 | pid[XT] u16 | scan scratch,
//      XT = TILE + FL*(lt-1) (the tile plus every digit's carry),
//      cf = carried tuples << 16 | tuples stored this tile
// NOUT: the pass writes 8-byte {key, row_id} tuples (16-byte input whose
// keys and row ids all fit 32 bits: narrow_ok and the histogram pass found no
// wider one), staged and carried as 8 bytes, at 8-byte positions of `out`.
// Both instances are launched; the one that does not match returns at once.
template <int WV, int NT, int ITEMS, bool NOUT>
__global__ __launch_bounds__(NT) void k_scatter(
    const hj_seg *__restrict__ segs, const uint32_t *__restrict__ chunk_seg, const hj_stats *__restrict__ st,
    uint64_t ch, uint32_t shift, uint32_t bits, uint32_t lt, const uint64_t *__restrict__ scan,
    uint4 *__restrict__ out, bool narrow_ok) {
	constexpr int TILE = NT * ITEMS;
	extern __shared__ uint4 lds_u4[];
	if (NOUT != (WV == 1 && narrow_ok && !(st->flags & HJ_DF_WIDE))) return;
	using tv = typename std::conditional<NOUT, uint2, uint4>::type;
	const uint32_t FL = 1u << bits;  // launch maximum: LDS layout
	const uint32_t CW = lt - 1;      // carry slots per digit
	const uint32_t XT = TILE + FL * CW;
	tv *tile = (tv *)lds_u4;
	tv *carry = tile + (size_t)XT * WV;
	uint64_t *cursor = (uint64_t *)(carry + (size_t)FL * CW * WV);
	uint32_t *tcnt = (uint32_t *)(cursor + FL);
	uint32_t *tbase = tcnt + FL;
	uint32_t *cf = tbase + FL;
	uint16_t *pid = (uint16_t *)(cf + FL);
	uint32_t *red = (uint32_t *)(pid + ((XT + 3) & ~1u));

	const uint32_t c = blockIdx.x;
	if (c >= st->total_chunks) return;
	const hj_seg g = segs[chunk_seg[c]];
	const uint32_t F = 1u << min(bits, g.bits), fmask = F - 1;
	const uint32_t cl = c - g.chunk_base;
	const uint64_t b = (uint64_t)cl * ch, e = min(g.len, b + ch);
	if (b >= e) return;  // whole workgroup: nothing to scatter
	for (uint32_t p = threadIdx.x; p < F; p += NT) {
		cursor[p] = scan[g.hist_base + (uint64_t)p * g.nchunks + cl];
		tcnt[p] = 0;
		cf[p] = 0;
	}
	__syncthreads();
	const uint32_t per = (F + NT - 1) / NT;  // digits per thread in the tile scan
	// Two register buffers: while one tile is sorted and written out, the
	// next one is in flight, and a buffer is refilled two tiles ahead right
	// after its tile's stores.  Every thread issues exactly ITEMS*WV loads per
	// refill (past the chunk end they re-read its last tuple) and both buffers
	// are processed on every trip, so a tile start waits only for its own
	// loads and the stores before them, never for the other buffer's loads.
	uint4 va[ITEMS][WV], vb[ITEMS][WV];
	auto load_tile = [&](uint4 (&v)[ITEMS][WV], uint64_t t0) {
#pragma unroll
		for (int k = 0; k < ITEMS; k++) {
			const uint64_t i = min(t0 + (uint64_t)(k * NT + threadIdx.x), e - 1);
#pragma unroll
			for (int w = 0; w < WV; w++) v[k][w] = hj_load_stream(g.in + i * WV + w);
		}
	};
	auto do_tile = [&](uint4 (&v)[ITEMS][WV], uint64_t t0) {
		const uint32_t n = t0 < e ? (uint32_t)min((uint64_t)TILE, e - t0) : 0u;
		uint32_t dig[ITEMS], rank[ITEMS];
#pragma unroll
		for (int k = 0; k < ITEMS; k++) {
			const uint32_t idx = k * NT + threadIdx.x;
			if (idx < n) {
				dig[k] = (uint32_t)(hj_key_of(v[k][0]) >> shift) & fmask;
				rank[k] = atomicAdd(&tcnt[dig[k]], 1u);
			}
		}
		__syncthreads();
		// digit p's stream this tile = its carry, then its tuples of the tile;
		// the first (cf[p] & 0xffff) tuples of it end on a line boundary and
		// are stored now
		uint32_t tot;
		{
			const uint32_t p0 = threadIdx.x * per;
			uint32_t sum = 0;
			for (uint32_t q = 0; q < per; q++)
				if (p0 + q < F) sum += (cf[p0 + q] >> 16) + tcnt[p0 + q];
			uint32_t run = hj_block_exscan<uint32_t>(sum, red, &tot);
			for (uint32_t q = 0; q < per; q++)
				if (p0 + q < F) {
					const uint32_t p = p0 + q, len = (cf[p] >> 16) + tcnt[p];
					tbase[p] = run;
					run += len;
					const uint64_t end = cursor[p] + len, aend = end - end % lt;
					cf[p] = (cf[p] & 0xffff0000u) | (aend > cursor[p] ? (uint32_t)(aend - cursor[p]) : 0u);
				}
		}
		__syncthreads();
		for (uint32_t x = threadIdx.x; x < F * CW; x += NT) {
			const uint32_t p = x / CW, j = x - p * CW;
			if (j < (cf[p] >> 16)) {
				const uint32_t q = tbase[p] + j;
#pragma unroll
				for (int w = 0; w < WV; w++) tile[q * WV + w] = carry[x * WV + w];
				pid[q] = (uint16_t)p;
			}
		}
#pragma unroll
		for (int k = 0; k < ITEMS; k++) {
			const uint32_t idx = k * NT + threadIdx.x;
			if (idx < n) {
				const uint32_t q = tbase[dig[k]] + (cf[dig[k]] >> 16) + rank[k];
				if constexpr (NOUT) {
					tile[q] = make_uint2(v[k][0].x, v[k][0].z);
				} else {
#pragma unroll
					for (int w = 0; w < WV; w++) tile[q * WV + w] = v[k][w];
				}
				pid[q] = (uint16_t)dig[k];
			}
		}
		__syncthreads();
		// consecutive lanes store consecutive words of a run
		for (uint32_t u = threadIdx.x; u < tot * WV; u += NT) {
			const uint32_t i = u / WV, w = u % WV, p = pid[i], j = i - tbase[p], fl = cf[p] & 0xffffu;
			if (j < fl) {
				if constexpr (NOUT) hj_store_stream8((uint2 *)out + cursor[p] + j, tile[u].x, tile[u].y);
				else hj_store_stream(&out[(cursor[p] + j) * WV + w], tile[u]);
			} else {
				carry[(p * CW + j - fl) * WV + w] = tile[u];
			}
		}
		load_tile(v, t0 + 2 * TILE);
		__syncthreads();
		for (uint32_t p = threadIdx.x; p < F; p += NT) {
			const uint32_t fl = cf[p] & 0xffffu;
			cursor[p] += fl;
			cf[p] = ((cf[p] >> 16) + tcnt[p] - fl) << 16;
			tcnt[p] = 0;
		}
		__syncthreads();
	};
	load_tile(va, b);
	load_tile(vb, b + TILE);
	for (uint64_t t0 = b; t0 < e; t0 += 2 * TILE) {
		do_tile(va, t0);
		do_tile(vb, t0 + TILE);  // an empty tile past the end changes nothing
	}
	// the chunk's last partial lines
	for (uint32_t x = threadIdx.x; x < F * CW * WV; x += NT) {
		const uint32_t s = x / WV, w = x % WV, p = s / CW, j = s - p * CW;
		if (j < (cf[p] >> 16)) {
			if constexpr (NOUT) hj_store_stream8((uint2 *)out + cursor[p] + j, carry[x].x, carry[x].y);
			else hj_store_stream(&out[(cursor[p] + j) * WV + w], carry[x]);
		}
	}
}

// ------------------------------------------------------------ planner
// pairs[q] and tiles[q] for every final partition q of one join.
//  passes 0: one pair (whole R, whole S) at offsets r_base / s_base.
//  passes 1: R digit q at pb[q], S digit q at pb[F1 + q].
//  passes 2: q = (p2 << b1) | p1; relation rel's partition at
//            pb[((rel*F1 + p1) * F2) + p2].
__global__ void k_plan(const uint64_t *__restrict__ pb, uint32_t passes, uint32_t Q, uint32_t b1, uint32_t F2,
                       uint64_t nr, uint64_t ns, uint64_t s_tile, hj_pair *__restrict__ pairs,
                       uint64_t *__restrict__ tiles) {
	const uint32_t F1 = 1u << b1;
	for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < Q; q += gridDim.x * blockDim.x) {
		hj_pair pr;
		if (passes == 0) {
			pr = hj_pair{0, nr, 0, ns};
		} else if (passes == 1) {
			pr = hj_pair{pb[q], pb[q + 1], pb[F1 + q], pb[F1 + q + 1]};
		} else {
			const uint32_t p1 = q & (F1 - 1), p2 = q >> b1;
			const uint64_t ir = ((uint64_t)p1 * F2) + p2, is = ((uint64_t)(F1 + p1) * F2) + p2;
			pr = hj_pair{pb[ir], pb[ir + 1], pb[is], pb[is + 1]};
		}
		pairs[q] = pr;
		const uint64_t rl = pr.r1 - pr.r0, sl = pr.s1 - pr.s0;
		tiles[q] = (rl && sl) ? (sl + s_tile - 1) / s_tile : 0;
	}
}

// ------------------------------------------------------------ join
// LDS (bc): keys u64[cap] | pays u64[cap] | heads u32[np2(cap)] | next u16[cap]
// LDS (lp): keys u64[cap] | pays u64[cap] | occ   u32[np2(cap)]
template <int WV, int NT, int TK, bool EMIT>
__global__ __launch_bounds__(NT) void k_join(const uint4 *__restrict__ Rb, const uint4 *__restrict__ Sb,
                                             const hj_pair *__restrict__ pairs,
                                             const uint64_t *__restrict__ tprefix, uint32_t npairs, uint32_t B,
                                             uint64_t s_tile, uint32_t cap, hj_stats *st, uint4 *pout,
                                             unsigned long long pcap, unsigned pchunk) {
	constexpr bool LP = TK == HJ_TK_LP;
	constexpr bool TWO = TK == HJ_TK_HR || TK == HJ_TK_HRO;  // count + place builds
	extern __shared__ uint64_t lds_u64[];
	__shared__ unsigned long long red[2 * (NT / HJ_WAVE)];
	__shared__ uint32_t sred[NT / HJ_WAVE + 1];
	__shared__ uint64_t s_item[4];
	__shared__ uint32_t s_pool;
	hj_tab<TK, WV> tab;
	tab.map(lds_u64, cap, &s_pool);
	const uint64_t total = tprefix[npairs];
	__shared__ unsigned long long emit_ws[EMIT ? 2 * (NT / HJ_WAVE) : 2];
	auto acc = hj_make_sink<EMIT>(pout, pcap, &st->pair_next, pchunk, emit_ws);
	bool dup = false, oversize = false;

	for (;;) {
		if (threadIdx.x == 0) {
			const uint64_t w = atomicAdd(&st->work_next, 1ULL);
			uint32_t lo = 0, hi = npairs;  // last pair with tprefix <= w
			while (hi - lo > 1) {
				const uint32_t mid = (lo + hi) >> 1;
				if (tprefix[mid] <= w) lo = mid; else hi = mid;
			}
			s_item[0] = w;
			s_item[1] = lo;
			s_item[2] = w < total ? w - tprefix[lo] : 0;
		}
		__syncthreads();
		const uint64_t w = s_item[0];
		const uint32_t q = (uint32_t)s_item[1];
		const uint64_t tix = s_item[2];
		__syncthreads();
		if (w >= total) break;
		const hj_pair pr = pairs[q];
		const uint64_t sb = pr.s0 + tix * s_tile, se = min(pr.s1, sb + s_tile);
		// A partition larger than one table is joined fill by fill.  For bc,
		// sc, hr and hro the pairs of the fills add up.  For lp (first match
		// only) that holds while no key is in two fills, so each lp fill
		// probes the later build tuples against its table and a hit flags a
		// duplicate key (the host then recomputes the join exactly); lp fills
		// of an oversized partition take half the table's slots, keeping
		// linear probing at load factor 1/2.
		const uint64_t fcap = LP && pr.r1 - pr.r0 > cap ? max(1u, hj_np2u(cap) / 2) : cap;
		for (uint64_t rc = pr.r0; rc < pr.r1; rc += fcap) {
			const uint32_t m = (uint32_t)min(fcap, pr.r1 - rc);
			tab.reset(m, NT);
			__syncthreads();
			// build tuples four per thread and step, loads issued together
			// (unconditional, clamped to the last tuple: a load under a per-lane
			// condition waits for its data before the next one issues)
			auto build = [&](auto &&op) {
				for (uint32_t i0 = threadIdx.x; i0 < m; i0 += 4 * NT) {
					uint4 t[4];
#pragma unroll
					for (int u = 0; u < 4; u++) t[u] = hj_load(&Rb[(rc + min(i0 + u * NT, m - 1)) * WV]);
#pragma unroll
					for (int u = 0; u < 4; u++)
						if (i0 + u * NT < m) op(i0 + u * NT, t[u]);
				}
			};
			if (TWO) {
				build([&](uint32_t, const uint4 &t) { tab.count(hj_key_of(t), B); });
				__syncthreads();
				tab.starts(sred, NT);
				__syncthreads();
				build([&](uint32_t, const uint4 &t) { tab.place(hj_key_of(t), hj_pay_of(t), B); });
			} else {
				build([&](uint32_t i, const uint4 &t) { tab.put(i, hj_key_of(t), hj_pay_of(t), B); });
			}
			__syncthreads();
			if (LP) {
				tab.lp_insert(m, B, NT, dup);
				__syncthreads();
				// keys of the later fills must not be in this one (the probe
				// tile's first item checks; every item sees the same fills)
				if (tix == 0)
					for (uint64_t i = rc + m + threadIdx.x; i < pr.r1; i += NT)
						if (tab.lp_has(hj_key_of(hj_load(&Rb[i * WV])), B)) dup = true;
			}
			// probe tuples U per thread and step; the next step's loads are
			// issued before the current step is probed
			constexpr int U = 4;
			auto load = [&](uint4 (&t)[U], uint64_t j0) {
#pragma unroll
				for (int u = 0; u < U; u++) t[u] = hj_load_stream(&Sb[min(j0 + (uint64_t)u * NT, se - 1) * WV]);
			};
			auto probe = [&](const uint4 (&t)[U], uint64_t j0) {
				uint64_t sk[U], sp[U];
				bool live[U];
#pragma unroll
				for (int u = 0; u < U; u++) {
					live[u] = j0 + (uint64_t)u * NT < se;
					sk[u] = hj_key_of(t[u]);
					sp[u] = hj_pay_of(t[u]);
				}
				tab.template probe<U>(sk, sp, live, B, acc);
			};
			// the same number of steps in every lane, loads clamped and never
			// under a lane condition, so each probe waits only for its own
			// buffer while the other one's loads are in flight
			if (sb < se) {
				const uint64_t step = (uint64_t)U * NT;
				const uint64_t nsteps = (se - sb + step - 1) / step;
				uint4 ta[U], tb[U];
				load(ta, sb + threadIdx.x);
				for (uint64_t k = 0; k < nsteps; k += 2) {
					const uint64_t j0 = sb + k * step + threadIdx.x;
					load(tb, j0 + step);
					probe(ta, j0);
					load(ta, j0 + 2 * step);
					probe(tb, j0 + step);  // past the end: every lane skips
				}
			}
			__syncthreads();
		}
	}
	acc.finish();
	const unsigned fl = (__any(dup) ? HJ_DF_DUPKEY : 0u) | (__any(oversize) ? HJ_DF_OVERSIZE : 0u);
	if (fl && (threadIdx.x & (HJ_WAVE - 1)) == 0) atomicOr(&st->flags, fl);
	hj_block_add2(acc.cnt, acc.sum, red, &st->matched, &st->checksum);
}

// ------------------------------------------------------------ pass 2
// Tile sort: every chunk (one LDS tile) of a segment is counting-sorted by
// one digit and written to the same range of `out`, so the stores of a chunk
// are one contiguous stream.  run[c*(F+1) + p] = start of digit p's run
// inside chunk c (u16), run[..F] = chunk length; cstart[c] = absolute
// position of chunk c.  Pass A of a two-pass join runs it over the input.
// The next tiles are loaded while a tile is sorted (two register buffers,
// as in k_scatter).
template <int WV, int NT, int ITEMS, bool BATCH = false>
__global__ __launch_bounds__(NT) void k_sortchunk(const hj_seg *__restrict__ segs, const uint32_t *__restrict__ cmap,
                                                  hj_stats *__restrict__ st, uint32_t shift, uint32_t bits,
                                                  uint4 *__restrict__ out, uint16_t *__restrict__ run,
                                                  uint64_t *__restrict__ cstart, bool narrow_ok, bool dense_ok,
                                                  uint8_t *__restrict__ tnarrow) {
	// BATCH: the segments are the relations of many joins, each with its own
	// digit width (hj_seg.bits; `bits` is the largest), `shift` is 0, tile
	// c's run table starts at segs[.].hist_base + (c - chunk_base) * (F + 1),
	// and cstart[c] receives the tile's address with its layout in the low
	// two bits (2: 4-byte positional words, 1: 8-byte, 0: 16-byte tuples)
	// instead of a position plus tnarrow[c].
	constexpr int TILE = NT * ITEMS;
	constexpr uint32_t LT = __builtin_ctz(TILE);  // bits of a tuple's index in its tile
	extern __shared__ uint4 lds_u4[];
	const uint32_t Fmax = 1u << bits;
	uint4 *tile = lds_u4;
	uint32_t *tcnt = (uint32_t *)(tile + TILE * WV);
	uint32_t *tbase = tcnt + Fmax;
	uint32_t *red = tbase + Fmax + 1;
	__shared__ uint32_t tflag;  // per tile: 1 a wide tuple, 2 a row id that is not position + 1
	const uint32_t total = (uint32_t)st->total_chunks;
	const uint32_t grid = gridDim.x;
	// tile c: source, output position, length (an empty tile past the end
	// reads `out`, whose first tuple always exists)
	// input tuples are WV words apart, or in_wv when only the heads of wider
	// tuples are read
	uint32_t sw = WV;
	uint64_t spos = 0;  // position of the tile's first tuple in its relation
	uint32_t tbits = bits;  // the tile's digit width
	uint64_t rbase = 0;     // first entry of the tile's run table
	auto tile_of = [&](uint32_t c, const uint4 *&src, uint64_t &pos, uint32_t &n) {
		if (c >= total) {
			src = out;
			pos = 0;
			n = 0;
			sw = WV;
			spos = 0;
			tbits = 0;
			rbase = 0;
			return;
		}
		const hj_seg &g = segs[cmap[c]];
		const uint64_t cl = c - g.chunk_base;
		sw = g.in_wv ? g.in_wv : WV;
		pos = g.out_start + cl * TILE;
		spos = cl * TILE;
		src = g.in + cl * TILE * sw;
		n = (uint32_t)min((uint64_t)TILE, g.len - cl * TILE);
		tbits = BATCH ? g.bits : bits;
		rbase = BATCH ? g.hist_base + cl * ((1ull << tbits) + 1) : (uint64_t)c * (Fmax + 1);
	};
	// every thread issues exactly ITEMS*WV loads (see k_scatter)
	auto load = [&](uint4 (&v)[ITEMS][WV], uint32_t c) {
		const uint4 *src;
		uint64_t pos;
		uint32_t n;
		tile_of(c, src, pos, n);
#pragma unroll
		for (int k = 0; k < ITEMS; k++) {
			const uint32_t idx = min((uint32_t)(k * NT) + threadIdx.x, n ? n - 1 : 0u);
#pragma unroll
			for (int w = 0; w < WV; w++) v[k][w] = hj_load_stream(src + (uint64_t)idx * sw + w);
		}
	};
	bool wide = false;
	unsigned long long narrow_cnt = 0, dense_cnt = 0;  // thread 0: tuples of this workgroup's 8- and 4-byte tiles
	auto do_tile = [&](uint4 (&v)[ITEMS][WV], uint32_t c) {
		const uint4 *src;
		uint64_t pos;
		uint32_t n;
		tile_of(c, src, pos, n);
		// this tile's digit width (tile_of is called again by the prefetch
		// below, so everything after it uses these copies)
		const uint32_t tb = tbits, F = 1u << tb, fmask = F - 1, per = (F + NT - 1) / NT;
		const uint64_t rb = rbase;
		// a positional word holds key >> tb above the LT index bits: keys
		// below 2^(32 - LT + tb) fit
		const uint32_t kshift = tb + 32 - LT;
		uint32_t dig[ITEMS], rank[ITEMS];
		bool tw = false, tp = false;
#pragma unroll
		for (int k = 0; k < ITEMS; k++) {
			const uint32_t idx = k * NT + threadIdx.x;
			if (idx < n) {
				dig[k] = (uint32_t)(hj_key_of(v[k][0]) >> shift) & fmask;
				rank[k] = atomicAdd(&tcnt[dig[k]], 1u);
				if (WV == 1) {
					tw |= (v[k][0].y | v[k][0].w) != 0;
					tp |= v[k][0].z != (uint32_t)(spos + idx + 1) || (kshift < 32 && (v[k][0].x >> kshift) != 0);
				}
			}
		}
		// a 16-byte tile whose keys and row ids all fit 32 bits is stored as
		// 8-byte {key, row_id} tuples in the first half of its range; one
		// whose row ids are also its relation positions + 1 (and whose keys
		// fit beside the index) as 4-byte {key >> bits, index} words in the
		// first quarter: pass B restores the row id from the position
		if (WV == 1) {
			const unsigned fw = (__any(tw) ? 1u : 0u) | (__any(tp) ? 2u : 0u);
			if (fw && (threadIdx.x & (HJ_WAVE - 1)) == 0) atomicOr(&tflag, fw);
		}
		__syncthreads();  // the counts and tflag are complete
		const uint32_t tf = WV == 1 ? tflag : 1u;
		const bool tile_narrow = WV == 1 && narrow_ok && !(tf & 1u);
		const bool tile_dense = tile_narrow && dense_ok && !(tf & 2u);
		wide |= tw;
		{
			const uint32_t p0 = threadIdx.x * per;
			uint32_t sum = 0;
			for (uint32_t q = 0; q < per; q++)
				if (p0 + q < F) sum += tcnt[p0 + q];
			uint32_t tot;
			uint32_t r = hj_block_exscan<uint32_t>(sum, red, &tot);
			if (n)
				for (uint32_t q = 0; q < per; q++)
					if (p0 + q < F) {
						tbase[p0 + q] = r;
						run[rb + p0 + q] = (uint16_t)r;
						r += tcnt[p0 + q];
					}
			if (threadIdx.x == 0 && n) {
				run[rb + F] = (uint16_t)n;
				const uint32_t code = tile_dense ? 2u : tile_narrow ? 1u : 0u;
				if (BATCH) cstart[c] = (uint64_t)(uintptr_t)(out + pos) | code;
				else cstart[c] = pos;
				if (WV == 1 && narrow_ok) {
					if (!BATCH) tnarrow[c] = (uint8_t)code;
					narrow_cnt += tile_narrow ? n : 0u;
					dense_cnt += tile_dense ? n : 0u;
				}
			}
		}
		__syncthreads();
		// a narrow tile is placed as 8-byte {key, row_id} words, so the store
		// below reads consecutive 8-byte LDS words (16-byte slots read half by
		// half hit the same banks from several lanes)
		uint2 *tile8 = (uint2 *)tile;
		uint32_t *tile4 = (uint32_t *)tile;
#pragma unroll
		for (int k = 0; k < ITEMS; k++) {
			const uint32_t idx = k * NT + threadIdx.x;
			if (idx < n) {
				const uint32_t q = tbase[dig[k]] + rank[k];
				if (WV == 1 && tile_dense) {
					tile4[q] = (v[k][0].x >> tb) << LT | idx;
				} else if (WV == 1 && tile_narrow) {
					tile8[q] = make_uint2(v[k][0].x, v[k][0].z);
				} else {
#pragma unroll
					for (int w = 0; w < WV; w++) tile[q * WV + w] = v[k][w];
				}
			}
		}
		__syncthreads();
		if (tile_dense) {
			uint32_t *dst4 = (uint32_t *)(out + pos);
			for (uint32_t u = threadIdx.x; u < n; u += NT) hj_store_stream4(&dst4[u], tile4[u]);
		} else if (tile_narrow) {
			uint2 *dst8 = (uint2 *)(out + pos);
			for (uint32_t u = threadIdx.x; u < n; u += NT) {
				const uint2 t = tile8[u];
				hj_store_stream8(&dst8[u], t.x, t.y);
			}
		} else {
			uint4 *dst = out + pos * WV;
			for (uint32_t u = threadIdx.x; u < n * WV; u += NT) hj_store_stream(&dst[u], tile[u]);
		}
		load(v, c + 2 * grid);
		for (uint32_t p = threadIdx.x; p < F; p += NT) tcnt[p] = 0;
		if (threadIdx.x == 0) tflag = 0;
		__syncthreads();
	};
	for (uint32_t p = threadIdx.x; p < Fmax; p += NT) tcnt[p] = 0;
	if (threadIdx.x == 0) tflag = 0;
	__syncthreads();
	uint4 va[ITEMS][WV], vb[ITEMS][WV];
	load(va, blockIdx.x);
	load(vb, blockIdx.x + grid);
	for (uint32_t c = blockIdx.x; c < total; c += 2 * grid) {
		do_tile(va, c);
		do_tile(vb, c + grid);  // an empty tile past the end changes nothing
	}
	// 16-byte tuples: flag a key or row_id above 32 bits (pass B then keeps
	// 16-byte tuples)
	if (WV == 1 && __any(wide) && (threadIdx.x & (HJ_WAVE - 1)) == 0) atomicOr(&st->flags, HJ_DF_WIDE);
	// one atomic per workgroup (one per tile on a single address serialises)
	if (WV == 1 && threadIdx.x == 0 && narrow_cnt) atomicAdd(&st->narrow_tuples, narrow_cnt);
	if (WV == 1 && threadIdx.x == 0 && dense_cnt) atomicAdd(&st->dense_tuples, dense_cnt);
}

// Two-pass joins partition without a histogram pass.  Pass A sorts every
// input tile by the first digit in place (k_sortchunk over the input
// segments); its run tables are the histogram: k_runs_to_hist turns them into
// the [segment][digit][tile] count matrix whose exclusive scan gives each
// tile's run of each digit its position in the fully partitioned order.
// Pass B (k_sortchunk_g) then cuts every first-level partition into tiles of
// that order, gathers each tile from the pass-A runs that cover it, and
// sorts it by the second digit, which gives the same layout as a
// histogram + scatter pass followed by the chunk-local second pass.

// Count matrix from pass-A run tables, as two kernels around a scan of
// per-group sums: the tiles of each segment form groups of 64 (segment 0's
// GR groups, then segment 1's GS), and part[seg][digit][group] (segment 1's
// rows after F*GR entries) = tuples of that digit in the group's tiles.
// After an exclusive scan of `part` in that order, k_runs_pos writes
// hist[g.hist_base + d*g.nchunks + cl] = the position of tile cl's run of
// digit d in the fully partitioned order (its group's scanned sum + the runs
// of the group's earlier tiles) and beg_t at the same index = where that run
// starts inside its tile.  Together that is the exclusive scan of the whole
// [segment][digit][tile] count matrix, with the scan's reduce phase done on
// the run tables (a third of the matrix's bytes).  A workgroup takes 64
// tiles x 64 digits; run[c*(F+1) + d] is u16, so a wave reads 64
// consecutive digits of one tile, and the matrix rows are written 64
// consecutive tiles at a time.
struct hj_grp {
	uint32_t seg, g, G, c0, nt, cl0;
};
__device__ __forceinline__ hj_grp hj_group_of(const hj_seg *segs, uint32_t grp, uint32_t GR, uint32_t GS) {
	hj_grp r;
	r.seg = grp < GR ? 0u : 1u;
	r.g = r.seg ? grp - GR : grp;
	r.G = r.seg ? GS : GR;
	const hj_seg &s = segs[r.seg];
	r.cl0 = r.g * 64;
	r.c0 = s.chunk_base + r.cl0;
	r.nt = min(64u, s.nchunks - r.cl0);
	return r;
}

__global__ __launch_bounds__(256) void k_runs_part(const uint16_t *__restrict__ run, const hj_seg *__restrict__ segs,
                                                   uint32_t F, uint32_t GR, uint32_t GS, uint32_t *__restrict__ part) {
	__shared__ uint32_t acc[4][64];
	if (blockIdx.x >= GR + GS) return;
	const hj_grp q = hj_group_of(segs, blockIdx.x, GR, GS);
	const uint32_t dj = threadIdx.x & 63, qt = threadIdx.x >> 6, d = blockIdx.y * 64 + dj;
	uint32_t sum = 0;
	if (d < F)
		for (uint32_t i = qt; i < q.nt; i += 4) {
			const uint16_t *r = run + (uint64_t)(q.c0 + i) * (F + 1) + d;
			sum += (uint32_t)r[1] - r[0];
		}
	acc[qt][dj] = sum;
	__syncthreads();
	if (threadIdx.x < 64 && d < F)
		part[(q.seg ? (uint64_t)F * GR : 0) + (uint64_t)d * q.G + q.g] = acc[0][dj] + acc[1][dj] + acc[2][dj] + acc[3][dj];
}

__global__ __launch_bounds__(256) void k_runs_pos(const uint16_t *__restrict__ run, const hj_seg *__restrict__ segs,
                                                  uint32_t F, uint32_t GR, uint32_t GS, const uint32_t *__restrict__ part,
                                                  uint32_t *__restrict__ hist, uint16_t *__restrict__ beg_t) {
	__shared__ uint32_t cnt[64][65];
	__shared__ uint32_t beg[64][65];
	__shared__ uint32_t acc[4][64];
	if (blockIdx.x >= GR + GS) return;
	const hj_grp q = hj_group_of(segs, blockIdx.x, GR, GS);
	const uint32_t d0 = blockIdx.y * 64;
	for (uint32_t e = threadIdx.x; e < 64 * 64; e += 256) {
		const uint32_t ci = e / 64, dj = e % 64, d = d0 + dj;
		uint32_t v = 0, a = 0;
		if (ci < q.nt && d < F) {
			const uint16_t *r = run + (uint64_t)(q.c0 + ci) * (F + 1) + d;
			a = r[0];
			v = (uint32_t)r[1] - a;
		}
		cnt[ci][dj] = v;
		beg[ci][dj] = a;
	}
	__syncthreads();
	// digit dj's runs over the group's tiles, 16 tiles per thread, turned
	// into positions in place
	const uint32_t dj = threadIdx.x & 63, qt = threadIdx.x >> 6, d = d0 + dj;
	uint32_t sum = 0;
	for (uint32_t i = 0; i < 16; i++) sum += cnt[qt * 16 + i][dj];
	acc[qt][dj] = sum;
	__syncthreads();
	if (d < F) {
		uint32_t run_pos = part[(q.seg ? (uint64_t)F * GR : 0) + (uint64_t)d * q.G + q.g];
		for (uint32_t k = 0; k < qt; k++) run_pos += acc[k][dj];
		for (uint32_t i = 0; i < 16; i++) {
			const uint32_t x = cnt[qt * 16 + i][dj];
			cnt[qt * 16 + i][dj] = run_pos;
			run_pos += x;
		}
	}
	__syncthreads();
	const hj_seg &g = segs[q.seg];
	for (uint32_t e = threadIdx.x; e < 64 * 64; e += 256) {
		const uint32_t dj2 = e / 64, ci = e % 64, d2 = d0 + dj2;
		if (ci < q.nt && d2 < F) {
			const uint64_t at = g.hist_base + (uint64_t)d2 * g.nchunks + q.cl0 + ci;
			hist[at] = cnt[ci][dj2];
			beg_t[at] = (uint16_t)beg[ci][dj2];
		}
	}
}

// The second pass's segments (one per first-level partition, s2 = seg*F1 +
// d1) and the join's per-first-digit work, by one workgroup of 1024
// threads: pb[s2] = first position of partition s2 (its group-0 sum in the
// scanned `part`, or the segment start when the segment has no tiles),
// segs2[s2] its bounds and pass-B tiles (ch tuples each, numbered
// consecutively over all s2), st->total_chunks their count, and for the
// join's items (d1, d2): tiles[d1] = probe tiles of cpt pass-B tiles each
// (0 when either side of d1 is empty) and dpre[d1] = items of all earlier
// first digits (F2 each), dpre[F1] = all.  It also ends the count matrix:
// hist[hist_len] = the total, the end of the last digit's row (k_runs_pos
// writes positions of runs, which start).  Dynamic LDS: 8 bytes per
// second-pass segment.
__global__ __launch_bounds__(1024) void k_segs2(const hj_seg *__restrict__ segs1, const uint32_t *__restrict__ part,
                                                uint32_t F1, uint32_t GR, uint32_t GS, uint64_t ch, const uint4 *in1,
                                                uint32_t wv, uint32_t b2, uint32_t cpt, hj_seg *__restrict__ segs2,
                                                uint64_t *__restrict__ pb, uint64_t *__restrict__ dpre,
                                                uint32_t *__restrict__ dtiles, hj_stats *__restrict__ st,
                                                uint32_t *__restrict__ hist, uint64_t hist_len) {
	extern __shared__ uint32_t s2lds[];
	uint32_t *slen = s2lds, *snch = s2lds + 2 * F1;
	__shared__ uint32_t red32[1024 / HJ_WAVE + 1];
	__shared__ uint64_t red64[1024 / HJ_WAVE + 1];
	const uint32_t nseg2 = 2 * F1, per = (nseg2 + 1023) / 1024;
	const uint64_t end = segs1[1].out_start + segs1[1].len;
	auto start = [&](uint32_t x) -> uint64_t {
		if (x >= nseg2) return end;
		const uint32_t seg = x / F1, d = x % F1;
		const hj_seg &g = segs1[seg];
		return g.nchunks ? part[(seg ? (uint64_t)F1 * GR : 0) + (uint64_t)d * (seg ? GS : GR)] : g.out_start;
	};
	const uint32_t x0 = threadIdx.x * per;
	uint32_t mine = 0;
	for (uint32_t k = 0; k < per; k++) {
		const uint32_t x = x0 + k;
		if (x < nseg2) {
			const uint64_t a = start(x), len = start(x + 1) - a;
			slen[x] = (uint32_t)len;
			snch[x] = (uint32_t)((len + ch - 1) / ch);
			mine += snch[x];
			pb[x] = a;
		}
	}
	uint32_t tot;
	uint32_t cb = hj_block_exscan<uint32_t>(mine, red32, &tot);
	for (uint32_t k = 0; k < per; k++) {
		const uint32_t x = x0 + k;
		if (x < nseg2) {
			hj_seg g{};
			g.out_start = pb[x];
			g.in = in1 + g.out_start * wv;
			g.len = slen[x];
			g.nchunks = snch[x];
			g.chunk_base = cb;
			g.hist_base = (uint64_t)cb << b2;
			g.bits = b2;
			segs2[x] = g;
			cb += g.nchunks;
		}
	}
	if (threadIdx.x == 0) {
		pb[nseg2] = end;
		st->total_chunks = tot;
		hist[hist_len] = (uint32_t)end;
	}
	__syncthreads();  // slen / snch complete
	const uint32_t perd = (F1 + 1023) / 1024, y0 = threadIdx.x * perd;
	uint64_t items = 0;
	for (uint32_t k = 0; k < perd; k++) {
		const uint32_t d1 = y0 + k;
		if (d1 < F1) {
			const uint32_t t = (slen[d1] && slen[F1 + d1]) ? (snch[F1 + d1] + cpt - 1) / cpt : 0u;
			dtiles[d1] = t;
			items += (uint64_t)t << b2;
		}
	}
	uint64_t tot64;
	uint64_t run = hj_block_exscan<uint64_t>(items, red64, &tot64);
	for (uint32_t k = 0; k < perd; k++) {
		const uint32_t d1 = y0 + k;
		if (d1 < F1) {
			dpre[d1] = run;
			run += (uint64_t)dtiles[d1] << b2;
		}
	}
	if (threadIdx.x == 0) dpre[F1] = tot64;
}

// One pass-B tile, everything k_sortchunk_g needs to stage and place it,
// so that a workgroup reads one uniform record per tile instead of a chain
// of dependent lookups (tile -> segment -> pass-A segment -> first run).
struct hj_gtile {
	uint64_t pos;         // first output position
	uint64_t at;          // offset of its digit's row in hist / beg_t
	uint64_t base;        // first tuple of its pass-A segment
	uint32_t c;           // tile (chunk) id: run table and chunk start
	uint32_t n;           // tuples
	uint32_t first;       // pass-A tile whose run holds the first tuple
	uint32_t nchunks;     // pass-A tiles of the segment
	uint32_t chunk_base;  // global id of the segment's first pass-A tile
	uint32_t ncov;        // pass-A runs (from `first` on) that start before the tile ends
};
// (48 bytes: pass B holds six records in scalar registers)
static_assert(sizeof(hj_gtile) == 48, "pass-B tile record size");

// gt[] = the pass-B tiles sorted by (segment, block of dblk first digits,
// tile index k, digit): tiles of neighbouring digits at about the same k read
// the same pass-A tiles, so processing them together lets them share those
// lines in L2.  first = the
// largest i with row[i] <= tile start (row = scanned counts of its digit).
// The tile counts and first tile ids of all 2 * F1 second-level segments
// are read into LDS once per workgroup for the rank and for finding a tile's
// segment (the last one whose first tile id is at most the tile's: empty
// segments share the next one's), with each digit block's tile count
// prefix, so that a tile's rank sums over its own block only (dynamic LDS:
// 24 * F1 bytes).
//
// The same kernel writes the join's work list, items in (first digit,
// second digit) order: item i is final partition q = (d2 << b1) | d1 with
// d1 = i >> b2 and d2 = i & (F2 - 1), so the items of one first-level
// partition are consecutive and the joins that read neighbouring runs of
// the same pass-B tiles run together.  Both sides of every (d1, d2) gather
// from the pass-B tiles of d1, so pairs[i] (the chunk id ranges of R's and
// S's segment d1) depends on d1 only, and so do its probe tiles (dtiles,
// from k_segs2): tpre[i] = dpre[d1] + d2 * dtiles[d1] is the exclusive scan
// of the items' tiles, tpre[Q] their total.
__global__ void k_first_run(const hj_seg *__restrict__ segs1, const hj_seg *__restrict__ segs2,
                            const hj_stats *__restrict__ st, uint32_t b1, uint32_t tile,
                            const uint32_t *__restrict__ hist, hj_gtile *__restrict__ gt, uint32_t b2,
                            const uint64_t *__restrict__ dpre, const uint32_t *__restrict__ dtiles,
                            hj_pair *__restrict__ pairs, uint64_t *__restrict__ tpre, uint32_t dblk) {
	extern __shared__ uint32_t nck[];
	const uint32_t total = (uint32_t)st->total_chunks, F1 = 1u << b1;
	uint32_t *cbs = nck + 2 * F1;
	// bpre[seg * nblk + j]: tiles of segment seg's digits below j * dblk
	uint32_t *bpre = cbs + 2 * F1;
	const uint32_t nblk = F1 / dblk;
	for (uint32_t i = threadIdx.x; i < 2 * F1; i += blockDim.x) {
		nck[i] = segs2[i].nchunks;
		cbs[i] = segs2[i].chunk_base;
	}
	__syncthreads();
	for (uint32_t i = threadIdx.x; i < 2 * nblk; i += blockDim.x) {
		const uint32_t seg = i / nblk, j = i % nblk;
		uint32_t a = 0;
		for (uint32_t d = 0; d < j * dblk; d++) a += nck[seg * F1 + d];
		bpre[i] = a;
	}
	__syncthreads();
	const uint32_t Q = F1 << b2;
	for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < Q; i += gridDim.x * blockDim.x) {
		const uint32_t d1 = i >> b2;
		pairs[i] = hj_pair{cbs[d1], (uint64_t)cbs[d1] + nck[d1], cbs[F1 + d1], (uint64_t)cbs[F1 + d1] + nck[F1 + d1]};
		tpre[i] = dpre[d1] + (uint64_t)(i & ((1u << b2) - 1)) * dtiles[d1];
	}
	if (blockIdx.x == 0 && threadIdx.x == 0) tpre[Q] = dpre[F1];
	for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < total; c += gridDim.x * blockDim.x) {
		uint32_t s2 = 0;
		for (uint32_t top = 2 * F1; top - s2 > 1;) {
			const uint32_t mid = (s2 + top) >> 1;
			if (cbs[mid] <= c) s2 = mid; else top = mid;
		}
		const hj_seg &g2 = segs2[s2];
		const uint32_t seg = s2 >> b1, d1 = s2 & (F1 - 1), k = c - g2.chunk_base;
		// pass-B order: blocks of dblk neighbouring first digits, and in a
		// block (tile index k, digit).  The tiles of digits d and d+1 that
		// read the same pass-A tiles (their indices differ by up to one, as
		// the partitions' runs accumulate slightly different lengths) then
		// lie within a few places of each other and run together on one XCD,
		// so a line that the two runs share is found in its L2 (C3: pass B
		// 1.23 -> 1.08 ms with blocks of 8 against all digits at each k)
		// segment 1 follows all of segment 0
		const uint32_t blk = d1 & ~(dblk - 1);
		uint32_t rank = (seg ? cbs[F1] : 0u) + bpre[seg * nblk + blk / dblk];
		const uint32_t *n_of = nck + seg * F1 + blk;
		for (uint32_t d = 0; d < dblk; d++) {
			const uint32_t n = n_of[d];
			rank += min(n, k) + (blk + d < d1 && n > k ? 1u : 0u);
		}
		const uint64_t ts = g2.out_start + (uint64_t)k * tile;
		const hj_seg &g1 = segs1[seg];
		const uint64_t at = g1.hist_base + (uint64_t)d1 * g1.nchunks;
		const uint32_t *row = hist + at;
		// largest lo with row[lo] <= ts (row[0] <= ts < row[nchunks]): start
		// at the proportional guess, gallop to a bracket, then bisect, so a
		// near-uniform digit takes a few dependent loads instead of log2
		const uint32_t nc = g1.nchunks;
		const uint64_t r0 = row[0], span = row[nc] - r0;
		uint32_t lo = 0, hi = nc;
		if (span) {
			const uint32_t g = (uint32_t)min((uint64_t)nc - 1, (ts - r0) * nc / span);
			uint32_t step = 1;
			if (row[g] <= ts) {
				lo = g;
				while (lo + step < hi && row[lo + step] <= ts) {
					lo += step;
					step <<= 1;
				}
				hi = min(hi, lo + step);
			} else {
				hi = g;
				while (hi - lo > step && row[hi - step] > ts) {
					hi -= step;
					step <<= 1;
				}
				lo = hi > step ? max(lo, hi - step) : lo;
			}
		}
		while (hi - lo > 1) {
			const uint32_t mid = (lo + hi) >> 1;
			if (row[mid] <= ts) lo = mid; else hi = mid;
		}
		hj_gtile t;
		t.pos = ts;
		t.at = at;
		t.base = g1.out_start;
		t.c = c;
		t.n = (uint32_t)min((uint64_t)tile, g2.out_start + g2.len - ts);
		t.first = lo;
		t.nchunks = g1.nchunks;
		t.chunk_base = g1.chunk_base;
		// covering runs: lo up to the last run that starts before the tile's
		// end (largest e in [lo, nc) with row[e] < ts + n), found by galloping
		// from lo and bisecting
		{
			const uint64_t te = ts + t.n;
			uint32_t a = lo, b = nc, step = 1;
			while (a + step < nc && row[a + step] < te) {
				a += step;
				step <<= 1;
			}
			b = min(nc, a + step);
			while (b - a > 1) {
				const uint32_t mid = (a + b) >> 1;
				if (row[mid] < te) a = mid; else b = mid;
			}
			t.ncov = t.n ? a - lo + 1 : 0;
		}
		gt[rank] = t;
	}
}

// Pass B: every tile of a first-level partition is gathered from the pass-A
// runs that cover it, sorted by the second digit in LDS and written to `out`
// at its place, with its run table written digit-major: run_t[p * rstride +
// c] = where digit p starts in tile c (p = F: the tile's length), the layout
// the join reads (a partition's runs over consecutive tiles are consecutive
// u16s; the tiles sharing a line are sorted at about the same time, so their
// 2-byte writes meet in L2).  Staging writes each tile
// position's source (offset in its segment) into LDS (two buffers); the
// next tile's loads are issued before the current tile is sorted, and every
// thread issues exactly ITEMS*WV of them (past the tile end they re-read its
// last tuple) so the waits stay exact.  row = scanned counts of one digit
// over the pass-A tiles of one segment (row[i+1] - row[i] = tile i's run
// length), brow = where those runs start inside their tiles, and pass-A tile
// i starts at tuple i*tileA of its segment.
//
// NIN: every pass-A tile was stored narrow (no tile was wide), so every
// source is one 8-byte tuple: 8-byte loads and an 8-byte LDS tile.  Both
// instances are launched for 16-byte tuples (which case holds is known on
// the device only) and the one that does not match returns at once.
template <int WV, int NT, int ITEMS, bool NIN, bool DENSE>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void k_sortchunk_g(
    const hj_gtile *__restrict__ gt, const uint32_t *__restrict__ hist, const uint16_t *__restrict__ beg_t,
    const hj_stats *__restrict__ st, uint32_t b1, uint32_t bits,
    const uint4 *__restrict__ in, uint4 *__restrict__ out, uint16_t *__restrict__ run_t, uint32_t rstride,
    uint64_t *__restrict__ cstart, bool narrow_ok, const uint8_t *__restrict__ tnarrow, uint32_t tileA, uint32_t grid2,
    uint32_t pmode) {
	constexpr int TILE = NT * ITEMS;
	extern __shared__ uint4 lds_u4[];
	const uint32_t F = 1u << bits, fmask = F - 1;
	// 16-byte tuples whose keys and row ids all fit 32 bits (pass A found no
	// wider one) leave as 8-byte {key, row_id}: half the writes here and
	// half the reads of the join.  Then every pass-A tile is narrow too, and
	// the NIN instance does the work.
	const bool narrow = WV == 1 && narrow_ok && !(st->flags & HJ_DF_WIDE);
	if (NIN != narrow) return;
	// pmode 1 / 2: this instance runs only when pass A did / did not write
	// positional words (a DENSE and a plain instance are both launched)
	if ((pmode == 1 && st->dense_tuples == 0) || (pmode == 2 && st->dense_tuples != 0)) return;
	using tv = typename std::conditional<NIN, uint2, uint4>::type;
	tv *tile = (tv *)lds_u4;
	// three source buffers: tile h is staged, loaded and sorted from
	// src[h % 3], so its sources are still there when it is sorted
	uint32_t *src0 = (uint32_t *)(tile + TILE * WV), *src1 = src0 + TILE, *src2 = src1 + TILE;
	uint32_t *tcnt = src2 + TILE;
	uint32_t *tbase = tcnt + F;
	uint32_t *red = tbase + F + 1;
	__shared__ uint32_t rbeg[NT], rsp[NT];  // long-run staging: run i's first tile position and source
	const uint32_t total = (uint32_t)st->total_chunks;
	const uint32_t per = (F + NT - 1) / NT;
	// When most tuples came as positional words, a tile gathers half the
	// bytes and three workgroups per CU keep their tiles in flight within an
	// XCD's L2; otherwise the first grid2 (two per CU) work and the rest
	// leave (workgroups are dispatched breadth-first over the CUs).
	uint32_t grid = gridDim.x;
	if (DENSE && 2 * st->dense_tuples < st->narrow_tuples) grid = min(grid, grid2);
	if (blockIdx.x >= grid) return;
	for (uint32_t p = threadIdx.x; p < F; p += NT) tcnt[p] = 0;

	// Staging is split in two so that its table reads are issued a tile
	// ahead: prefetch() reads, for thread i, the run of pass-A tile first + i
	// (every thread the same loads, at clamped indices), and stage() writes
	// the source offset of every tuple of tile t into srcb[0..n) from those
	// values; bit 31 marks a pass-A tile stored as 8-byte tuples (segments
	// hold < 2^31 tuples then).  A tile covering more than NT pass-A runs
	// reads the rest in stage() itself.
	struct pre_t {
		uint32_t h0, h1, b, nar;
	};
	// A staged source is its offset in the pass-A segment; with narrow_ok,
	// bit 31 marks a pass-A tile stored in 8 or 4 bytes per tuple and, in the
	// DENSE instances (launched when pass A may have written positional
	// words: segments below 2^30), bit 30 one of 4-byte positional words.
	auto src_code = [](uint32_t nar) -> uint32_t { return nar == 2 ? 3u << 30 : 1u << 31; };
	const uint32_t omask = DENSE ? 0x3fffffffu : 0x7fffffffu;
	const uint32_t LTA = __ffs(tileA) - 1;  // index bits of a positional word
	// (no branch around the loads: a load whose result merges with another
	// value makes the compiler wait for every load in flight; an empty
	// record has nchunks = 1 and reads entry 0 of the tables)
	// Thread i reads the entries of covering run first + i; threads past the
	// covering runs read the last one's again (same lines: only the tables of
	// the runs the tile needs are fetched).
	auto prefetch = [&](const hj_gtile &t) -> pre_t {
		pre_t p;
		const uint32_t *row = hist + t.at;
		const uint32_t ii = t.first + min(threadIdx.x, max(t.ncov, 1u) - 1);
		p.h0 = row[ii];
		p.h1 = row[ii + 1];
		p.b = beg_t[t.at + ii];
		p.nar = tnarrow[t.chunk_base + ii];
		return p;
	};
	auto run_src = [&](uint32_t *srcb, uint64_t pos, uint32_t n, uint32_t ii, uint64_t H, uint64_t H1, uint32_t b,
	                   uint32_t nar) {
		const uint64_t lo = max(H, pos), hi = min(H1, pos + n);
		uint32_t sp = ii * tileA + b + (uint32_t)(lo - H);
		if (WV == 1 && narrow_ok && lo < hi && nar) sp |= src_code(nar);  // nar is unset unless narrow_ok
		for (uint64_t q = lo; q < hi; q++) srcb[q - pos] = sp + (uint32_t)(q - lo);
	};
	// A thread writes its own run's sources, so a tile made of a few long
	// runs (a heavy digit under skew: one run can fill the whole tile) would
	// be staged by a few threads in a loop of up to TILE steps while the rest
	// wait.  For a tile covered by at most 32 runs, the runs' tile offsets and
	// sources go to LDS instead and every thread stages its ITEMS positions,
	// finding each one's run by bisection (run offsets are non-decreasing in
	// the run).
	auto stage = [&](const pre_t &p, const hj_gtile &t, uint32_t *srcb) {
		const uint64_t pos = t.pos;
		const uint32_t n = t.n;
		if (n == 0) return;
		const bool mine = threadIdx.x < t.ncov;
		if (t.ncov <= 32) {  // uniform
			const uint64_t lo = max(p.h0, pos);
			uint32_t sp = (t.first + threadIdx.x) * tileA + p.b + (uint32_t)(lo - p.h0);
			if (WV == 1 && narrow_ok && p.nar) sp |= src_code(p.nar);
			rbeg[threadIdx.x] = mine ? (uint32_t)(min(max(p.h0, pos), pos + n) - pos) : n;
			rsp[threadIdx.x] = sp;
			__syncthreads();
			{
				const uint32_t nr = t.ncov;
#pragma unroll
				for (int k = 0; k < ITEMS; k++) {
					const uint32_t f = k * NT + threadIdx.x;
					if (f >= n) continue;
					uint32_t a = 0, b = nr;  // largest run a with rbeg[a] <= f
					while (b - a > 1) {
						const uint32_t mid = (a + b) >> 1;
						if (rbeg[mid] <= f) a = mid; else b = mid;
					}
					srcb[f] = rsp[a] + (f - rbeg[a]);
				}
				return;
			}
		}
		if (mine) run_src(srcb, pos, n, t.first + threadIdx.x, p.h0, p.h1, p.b, p.nar);
		if (t.ncov <= NT) return;  // uniform
		const uint32_t *row = hist + t.at;
		const uint16_t *brow = beg_t + t.at;
		for (uint32_t i = NT; i < t.ncov; i += NT) {
			const uint32_t ii = t.first + i + threadIdx.x;
			if (i + threadIdx.x < t.ncov)
				run_src(srcb, pos, n, ii, row[ii], row[ii + 1], brow[ii], tnarrow[t.chunk_base + ii]);
		}
	};
	// Load tile t into v.  For 16-byte tuples, sel gets two bits per item
	// (pass-A tile stored narrow, upper half of the loaded pair) and v the
	// raw 16-byte load; unpack() picks the tuple out when the tile is sorted,
	// since picking it here would wait for every load before the next issues.
	// A source e: offset sp = e & omask in the pass-A segment, pass-A tile
	// start t0 = sp & ~(tileA - 1).  A tuple at offset o of that tile lies at
	// 8-byte index 2*t0 + o = sp + t0 when stored narrow and at 4-byte index
	// 4*t0 + o = sp + 3*t0 when stored as a positional word; the load of the
	// aligned 8 (16) bytes holding it keeps one load per tuple in every case,
	// and unpack() reads e again to pick the tuple out (picking it at load
	// time would wait for every load before the next issues).  32-bit index
	// arithmetic: offsets are below 2^31, and below 2^30 where positional
	// words occur.
	auto load = [&](tv (&v)[ITEMS][WV], const hj_gtile &t, const uint32_t *srcb) {
		const uint32_t n = t.n;
		// source offsets are relative to the pass-A segment's first tuple
		const uint4 *base = in + t.base * WV;
#pragma unroll
		for (int k = 0; k < ITEMS; k++) {
			const uint32_t f = min((uint32_t)(k * NT) + threadIdx.x, n ? n - 1 : 0u);
			const uint32_t e = n ? srcb[f] : 0;
			// cached loads: a run's neighbours belong to tiles of nearby digits
			// that the same XCD works on at the same time
			if constexpr (NIN) {
				const uint32_t sp = e & omask, t0 = sp & ~(tileA - 1);
				uint32_t i8 = sp + t0;
				if constexpr (DENSE) {
					const uint32_t db = (e >> 30) & 1u;
					i8 += (((sp + 3u * t0) >> 1) - i8) & (0u - db);
				}
				v[k][0] = hj_load8((const uint2 *)base + i8);
			} else if (WV == 1) {
				const uint32_t nw = narrow_ok ? e >> 31 : 0u;
				const uint32_t sp = narrow_ok ? (e & omask) : e, t0 = sp & ~(tileA - 1);
				uint32_t i16 = sp + ((((sp + t0) >> 1) - sp) & (0u - nw));
				if constexpr (DENSE) {
					const uint32_t db = (e >> 30) & nw;
					i16 += (((sp + 3u * t0) >> 2) - i16) & (0u - db);
				}
				v[k][0] = hj_load(base + i16);
			} else {
				const uint64_t sp = e;
#pragma unroll
				for (int w = 0; w < WV; w++) v[k][w] = hj_load(base + sp * WV + w);
			}
		}
	};
	// Pass B leaves every key shifted right by the first digit's b1 bits
	// (key >> b1): all keys of a first-level partition share those bits, so
	// equality and the join's bucket (key >> B = (key >> b1) >> (B - b1)) are
	// unchanged, and a positional word w holds exactly that, w >> LTA; its
	// row id is its pass-A tile start + 1 + (w & (tileA - 1)).
	auto unpack = [&](tv (&v)[ITEMS][WV], const hj_gtile &t, const uint32_t *srcb) {
		const uint32_t n = t.n;
#pragma unroll
		for (int k = 0; k < ITEMS; k++) {
			if constexpr (NIN) {
				if constexpr (DENSE) {
					const uint32_t e = n ? srcb[min((uint32_t)(k * NT) + threadIdx.x, n ? n - 1 : 0u)] : 0;
					const uint32_t sp = e & omask, t0 = sp & ~(tileA - 1);
					// the word's 8-byte pair: index sp + 3*t0, whose parity is sp's
					const uint32_t w = (sp & 1) ? v[k][0].y : v[k][0].x;
					const bool dn = (e >> 30) & 1u;
					v[k][0].y = dn ? t0 + 1 + (w & (tileA - 1)) : v[k][0].y;
					v[k][0].x = dn ? w >> LTA : v[k][0].x >> b1;
				} else {
					v[k][0].x >>= b1;
				}
			} else if constexpr (WV == 1) {
				const uint32_t e = n ? srcb[min((uint32_t)(k * NT) + threadIdx.x, n ? n - 1 : 0u)] : 0;
				const uint32_t nw = narrow_ok ? e >> 31 : 0u;
				const uint32_t sp = narrow_ok ? (e & omask) : e, t0 = sp & ~(tileA - 1);
				const uint4 h = v[k][0];
				uint4 o;
				if (DENSE && nw && ((e >> 30) & 1u)) {
					const uint32_t q = (sp + 3u * t0) & 3;
					const uint32_t w = q == 0 ? h.x : q == 1 ? h.y : q == 2 ? h.z : h.w;
					o = make_uint4(w >> LTA, 0u, t0 + 1 + (w & (tileA - 1)), 0u);
				} else if (nw) {
					o = ((sp + t0) & 1) ? make_uint4(h.z >> b1, 0u, h.w, 0u) : make_uint4(h.x >> b1, 0u, h.y, 0u);
				} else {
					const uint64_t key = hj_key_of(h) >> b1;
					o = make_uint4((uint32_t)key, (uint32_t)(key >> 32), h.z, h.w);
				}
				v[k][0] = o;
			} else {
				const uint64_t key = hj_key_of(v[k][0]) >> b1;
				v[k][0].x = (uint32_t)key;
				v[k][0].y = (uint32_t)(key >> 32);
			}
		}
	};
	auto sort_store = [&](tv (&v)[ITEMS][WV], const hj_gtile &t, const uint32_t *srcb) {
		const uint64_t pos = t.pos;
		const uint32_t n = t.n, c = t.c;
		unpack(v, t, srcb);
		uint32_t dig[ITEMS], rank[ITEMS];
#pragma unroll
		for (int k = 0; k < ITEMS; k++) {
			const uint32_t idx = k * NT + threadIdx.x;
			if (idx < n) {
				// keys arrive shifted by b1 (unpack)
				dig[k] = v[k][0].x & fmask;
				rank[k] = atomicAdd(&tcnt[dig[k]], 1u);
			}
		}
		__syncthreads();
		// Every thread clears the counters it owns once it has read them (no
		// other thread reads them), so the next tile's counting, which has no
		// barrier of its own before it, finds them zero: the barriers below
		// order the clearing before it.
		if (per == 1) {
			// one digit per thread (F <= NT), without a loop: a loop that
			// stores makes the compiler wait for every load in flight first
			const uint32_t cnt = threadIdx.x < F ? tcnt[threadIdx.x] : 0u;
			if (threadIdx.x < F) tcnt[threadIdx.x] = 0;
			uint32_t tot;
			const uint32_t r = hj_block_exscan<uint32_t>(cnt, red, &tot);
			if (n && threadIdx.x < F) {
				tbase[threadIdx.x] = r;
				run_t[(uint64_t)threadIdx.x * rstride + c] = (uint16_t)r;
			}
			if (threadIdx.x == 0 && n) {
				run_t[(uint64_t)F * rstride + c] = (uint16_t)n;
				cstart[c] = pos;
			}
		} else {
			const uint32_t p0 = threadIdx.x * per;
			uint32_t sum = 0;
			for (uint32_t q = 0; q < per; q++)
				if (p0 + q < F) sum += tcnt[p0 + q];
			uint32_t tot;
			uint32_t r = hj_block_exscan<uint32_t>(sum, red, &tot);
			for (uint32_t q = 0; q < per; q++)
				if (p0 + q < F) {
					if (n) {
						tbase[p0 + q] = r;
						run_t[(uint64_t)(p0 + q) * rstride + c] = (uint16_t)r;
					}
					r += tcnt[p0 + q];
					tcnt[p0 + q] = 0;
				}
			if (threadIdx.x == 0 && n) {
				run_t[(uint64_t)F * rstride + c] = (uint16_t)n;
				cstart[c] = pos;
			}
		}
		__syncthreads();
#pragma unroll
		for (int k = 0; k < ITEMS; k++) {
			const uint32_t idx = k * NT + threadIdx.x;
			if (idx < n) {
				const uint32_t q = tbase[dig[k]] + rank[k];
#pragma unroll
				for (int w = 0; w < WV; w++) tile[q * WV + w] = v[k][w];
			}
		}
		__syncthreads();
		// (the other instance returns when the output is narrow, so only this
		// one writes 8-byte tuples)
		if constexpr (NIN) {
			uint2 *dst8 = (uint2 *)out + pos;
#pragma unroll
			for (int k = 0; k < ITEMS; k++) {  // fixed trip count: no loop to wait before
				const uint32_t u = k * NT + threadIdx.x;
				if (u < n) hj_store_stream8(&dst8[u], tile[u].x, tile[u].y);
			}
		} else {
			uint4 *dst = out + pos * WV;
#pragma unroll
			for (int k = 0; k < ITEMS * WV; k++) {
				const uint32_t u = k * NT + threadIdx.x;
				if (u < n * WV) hj_store_stream(&dst[u], tile[u]);
			}
		}
	};

	// Step j of this workgroup takes ordered tile j*grid + xcd*(grid/8) + lane
	// where workgroups are dispatched round-robin over the 8 XCDs (xcd =
	// blockIdx % 8): at each step an XCD works on a contiguous run of the
	// order, i.e. neighbouring digits at about the same k, in its own L2.  grid is
	// a multiple of 8; past the end a step takes an empty tile.
	const uint32_t lane = blockIdx.x >> 3, xslot = (blockIdx.x & 7) * (grid >> 3);
	auto tile_at = [&](uint32_t j) -> hj_gtile {
		const uint32_t w = j * grid + xslot + lane;
		if (w < total) return gt[w];
		hj_gtile e{};  // empty: n = 0, reads the first tuple of `in`
		e.nchunks = 1;
		return e;
	};
	if (total == 0) return;
	// Half step h (tile h is in flight into v[h&1], tile h+1 is staged in
	// src[(h+1)%3], tile h+2's tables are prefetched): prefetch tile h+3's
	// tables, load tile h+1, stage tile h+2 into src[(h+2)%3] (free: its last
	// reader, the sort of tile h-1, is behind that sort's barriers), then
	// sort tile h, which reads its sources in src[h%3] again.  Stores count
	// in the same hardware counter as loads, so this order is the one in
	// which every wait is for exactly the oldest loads it needs: the staging
	// waits for a prefetch issued a half step earlier, the sort for a load
	// issued a half step earlier, and neither for the loads issued after
	// them.  Records are read two half steps before their prefetch.
	tv va[ITEMS][WV], vb[ITEMS][WV];
	hj_gtile ta = tile_at(0), tb = tile_at(1), tc = tile_at(2), td = tile_at(3);
	stage(prefetch(ta), ta, src0);
	const pre_t p1 = prefetch(tb);
	pre_t p2 = prefetch(tc), p3;
	__syncthreads();
	load(va, ta, src0);
	stage(p1, tb, src1);
	__syncthreads();
	for (uint32_t j = 0; j * grid < total; j += 2) {
		// ta, tb in src0, src1; src2 free
		const hj_gtile te = tile_at(j + 4);
		p3 = prefetch(td);
		load(vb, tb, src1);
		stage(p2, tc, src2);
		sort_store(va, ta, src0);
		const hj_gtile tf = tile_at(j + 5);
		p2 = prefetch(te);
		load(va, tc, src2);
		stage(p3, td, src0);
		sort_store(vb, tb, src1);  // an empty tile past the end changes nothing
		// tc, td in src2, src0; src1 free
		uint32_t *const s0 = src0;
		src0 = src2;
		src2 = src1;
		src1 = s0;
		ta = tc;
		tb = td;
		tc = te;
		td = tf;
	}
}

// key and payload of a loaded tuple: a 16-byte tuple's head, or a narrow
// {u32 key, u32 row_id} pair
__device__ __forceinline__ uint64_t key_of(const uint4 &t) { return hj_key_of(t); }
__device__ __forceinline__ uint64_t pay_of(const uint4 &t) { return hj_pay_of(t); }
__device__ __forceinline__ uint64_t key_of(const uint2 &t) { return t.x; }
__device__ __forceinline__ uint64_t pay_of(const uint2 &t) { return t.y; }

// Work-item claim of the two-pass joins (thread 0 of a workgroup).  Items
// [tprefix[i0], tprefix[i1]) of pairs [i0, i1) form queue c (i0 = npairs*c/8,
// i1 = npairs*(c+1)/8); xq[c] counts the positions taken.  The queue bounds
// and the position are read together; past the end, the next queues are
// tried.  The item's pair: largest a in [i0, i1) with tprefix[a] <= w, found
// from the proportional guess (exact for near-uniform partitions: one pair of
// independent loads) and bisection of the side it leaves.  (A position taken
// ahead, while the previous item runs, would hold that item back behind a
// heavy one under skew: C3 at theta 1.5, join 0.85 -> 1.12 ms.)  Returns w
// (total when every queue is empty), the pair and the item's tile within it.
__device__ __forceinline__ void hj_claim_item(const uint64_t *__restrict__ tprefix, uint32_t npairs, uint64_t total,
                                              unsigned long long *xq, uint32_t &cur, uint64_t &w_out,
                                              uint32_t &pair_out, uint64_t &tix_out) {
	uint64_t w = total, tix = 0;
	uint32_t pair = 0;
	for (uint32_t tried = 0; tried < 8; tried++, cur = (cur + 1) & 7) {
		const uint32_t i0 = (uint32_t)((uint64_t)npairs * cur / 8), i1 = (uint32_t)((uint64_t)npairs * (cur + 1) / 8);
		const uint64_t w0 = tprefix[i0], w1 = tprefix[i1];
		const uint64_t k = atomicAdd(&xq[cur], 1ULL);
		if (w0 + k < w1) {
			w = w0 + k;
			const uint32_t g = i0 + (uint32_t)((w - w0) * (uint64_t)(i1 - i0) / (w1 - w0));
			const uint64_t tg = tprefix[g], tg1 = tprefix[g + 1];
			uint32_t a = i0, b = i1;  // tprefix[a] <= w < tprefix[b]
			uint64_t ta = w0;
			if (tg <= w) {
				a = g;
				ta = tg;
				if (w < tg1) b = g + 1;
			} else {
				b = g;
			}
			while (b - a > 1) {
				const uint32_t mid = (a + b) >> 1;
				const uint64_t tm = tprefix[mid];
				if (tm <= w) {
					a = mid;
					ta = tm;
				} else {
					b = mid;
				}
			}
			pair = a;
			tix = w - ta;
			break;
		}
	}
	w_out = w;
	pair_out = pair;
	tix_out = tix;
}

// Join over chunk runs.  A work item is (final partition, probe tile = a range
// of pass-B chunks of S), items in k_plan2's (first digit, second digit)
// order.  The items are cut into eight queues of consecutive first-level
// partitions, one per XCD (workgroups are dispatched round-robin over the
// XCDs, so blockIdx % 8 names a workgroup's XCD): the 64 workgroups of an XCD
// then join the 32 second digits of one or two first-level partitions at a
// time, and a 128-byte line that two neighbouring runs of a pass-B tile share
// is read from HBM once and found in that XCD's L2 the second time.  A
// workgroup whose queue is empty takes items from the next queues, and leaves
// once it has found all eight empty (the cursors only grow).  Descriptors of up to NT
// chunk runs (absolute start, flat offset) are staged in LDS; the staged runs
// form one flat sequence of tuples that the waves split into contiguous
// ranges, so the lanes of a wave read consecutive tuples whatever the run
// lengths.  Build side: runs are copied into the LDS table until it holds
// `cap` tuples, and a run may be split between two fills (bc joins every fill
// against the whole probe tile; lp flags a partition that needs more than one
// fill, and the host recomputes that join exactly).  Probe side: each wave
// loads the next RU tuples per lane while it probes the current ones.
// LDS: keys u64[cap] | pays u64[cap] | heads u32[np2(cap)] | next u16[cap]
template <int WV, int NT, int TK, bool EMIT, bool NARROW>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void k_join2(const uint4 *__restrict__ P, const uint64_t *__restrict__ cstart,
                                              const uint16_t *__restrict__ run_t, uint32_t stride, uint32_t b1,
                                              const hj_pair *__restrict__ pairs,
                                              const uint64_t *__restrict__ tprefix, uint32_t npairs, uint32_t B,
                                              uint32_t cpt, uint32_t cap, hj_stats *st, uint4 *pout,
                                              unsigned long long pcap, unsigned pchunk, bool narrow_ok) {
	constexpr bool LP = TK == HJ_TK_LP;
	// NARROW reads the 8-byte {key, row_id} tuples that pass B writes when
	// every key and row id fits 32 bits (see k_sortchunk_g).  Which layout
	// pass B chose is known on the device only, so both instances are
	// launched and the one that does not match returns at once.
	static_assert(!NARROW || WV == 1, "narrow tuples are 16-byte tuples packed");
	// pass B leaves keys shifted right by b1 bits: bucket (key >> B) is then
	// (key >> b1) >> hs
	const uint32_t hs = B - b1;
	if (NARROW != (WV == 1 && narrow_ok && !(st->flags & HJ_DF_WIDE))) return;
	constexpr bool TWO = TK == HJ_TK_HR || TK == HJ_TK_HRO;  // count + place builds
	constexpr uint32_t NW = NT / HJ_WAVE;
	constexpr uint32_t RU = 6;                  // tuples per lane per step
	constexpr uint32_t STEP = HJ_WAVE * RU;     // tuples per wave per step
	extern __shared__ uint64_t lds_u64[];
	__shared__ unsigned long long red[2 * NW];
	__shared__ uint32_t sred[NW + 1];
	__shared__ uint64_t s_item[4];
	__shared__ uint64_t r_src[NT];      // absolute position of a staged run's first tuple
	__shared__ uint32_t r_off[NT + 1];  // flat offset of a staged run; r_off[NT] = total
	__shared__ uint32_t s_pool;
	hj_tab<TK, WV> tab;
	tab.map(lds_u64, cap, &s_pool);
	const uint32_t lane = threadIdx.x & (HJ_WAVE - 1), wid = threadIdx.x / HJ_WAVE;
	const uint64_t total = tprefix[npairs];
	const uint32_t b2 = B - b1, m2 = (1u << b2) - 1;
	uint32_t cur = blockIdx.x & 7;  // this workgroup's queue (its XCD)
	__shared__ unsigned long long emit_ws[EMIT ? 2 * (NT / HJ_WAVE) : 2];
	auto acc = hj_make_sink<EMIT>(pout, pcap, &st->pair_next, pchunk, emit_ws);
	bool dup = false, oversize = false;

	// Stage the digit-p2 runs of chunks [c0, c1), c1 - c0 <= NT; the first run
	// starts `skip` tuples in.  Returns the flat total; the caller syncs
	// before reading r_src / r_off.
	auto stage = [&](uint64_t c0, uint64_t c1, uint32_t p2, uint32_t skip) -> uint32_t {
		const uint64_t c = c0 + threadIdx.x;
		uint32_t len = 0;
		uint64_t src = 0;
		if (c < c1) {
			uint32_t a = run_t[(uint64_t)p2 * stride + c];
			len = run_t[(uint64_t)(p2 + 1) * stride + c] - a;
			if (threadIdx.x == 0) {
				a += skip;
				len -= skip;
			}
			src = cstart[c] + a;
		}
		uint32_t tot;
		const uint32_t ex = hj_block_exscan<uint32_t>(len, sred, &tot);
		r_src[threadIdx.x] = src;
		r_off[threadIdx.x] = ex;
		if (threadIdx.x == NT - 1) r_off[NT] = tot;
		return tot;
	};
	// staged run holding flat index f (largest k with r_off[k] <= f)
	auto find_run = [&](uint32_t f) -> uint32_t {
		uint32_t lo = 0, hi = NT;
		while (hi - lo > 1) {
			const uint32_t mid = (lo + hi) >> 1;
			if (r_off[mid] <= f) lo = mid; else hi = mid;
		}
		return lo;
	};
	// A loaded tuple: narrow ones as the {key, row_id} pair the load fills.
	using tup_t = typename std::conditional<NARROW, uint2, uint4>::type;
	// Load the step of flat indices f0 + u*64 + lane (f0 < fe); k only moves
	// forward.  Every lane loads RU tuples, those past fe the tuple at fe - 1
	// (the callers skip them): a load under a per-lane condition must be
	// merged with the register's old value, and for narrow pairs that merge
	// waits for the data, leaving one load in flight instead of RU.  Cached
	// loads: the line two neighbouring runs share stays in L2 for the second
	// reader (streaming loads fetched it twice).
	auto load_step = [&](uint32_t f0, uint32_t fe, uint32_t &k, tup_t (&t)[RU]) {
#pragma unroll
		for (uint32_t u = 0; u < RU; u++) {
			const uint32_t f = min(f0 + u * HJ_WAVE + lane, fe - 1);
			while (r_off[k + 1] <= f) k++;
			const uint64_t at = r_src[k] + (f - r_off[k]);
			if constexpr (NARROW) t[u] = *((const uint2 *)P + at);
			else t[u] = P[at * WV];
		}
	};
	auto probe_step = [&](uint32_t f0, uint32_t fe, const tup_t (&t)[RU]) {
		uint64_t kk[RU], pp[RU];
		bool live[RU];
#pragma unroll
		for (uint32_t u = 0; u < RU; u++) {
			live[u] = f0 + u * HJ_WAVE + lane < fe;
			kk[u] = key_of(t[u]);
			pp[u] = pay_of(t[u]);
		}
		tab.template probe<(int)RU>(kk, pp, live, hs, acc);
	};

	for (;;) {
		if (threadIdx.x == 0) {
			uint64_t w, tix;
			uint32_t lo;
			hj_claim_item(tprefix, npairs, total, st->xq, cur, w, lo, tix);
			s_item[0] = w;
			s_item[1] = lo;
			s_item[2] = tix;
		}
		__syncthreads();
		const uint64_t w = s_item[0];
		const uint32_t q = (uint32_t)s_item[1];
		const uint64_t tix = s_item[2];
		__syncthreads();
		if (w >= total) break;
		const hj_pair pr = pairs[q];
		const uint32_t p2 = q & m2;  // item q's second digit
		const uint64_t sb = pr.s0 + tix * cpt, se = min(pr.s1, sb + cpt);
		uint64_t rc = pr.r0;
		uint32_t roff = 0;  // tuples of run rc already taken by an earlier fill
		// Gather the next fill of up to `cap` build tuples from (rc, roff) on,
		// advancing (rc, roff, m).  mode 0 inserts (bc, sc, lp), 1 counts and
		// 2 places (hr, hro: two gathers over the same runs).
		auto fill = [&](int mode, uint64_t &frc, uint32_t &froff, uint32_t &m) {
			for (;;) {
				const uint64_t c1 = min(pr.r1, frc + NT);
				const uint32_t n = (uint32_t)(c1 - frc);
				const uint32_t tot = stage(frc, c1, p2, froff);
				__syncthreads();
				const uint32_t T = min(tot, cap - m);
				// runs taken whole form a prefix; the next one may be taken in part
				const uint32_t nfull = __syncthreads_count(threadIdx.x < n && r_off[threadIdx.x + 1] <= T);
				const uint32_t roff_new = nfull < n ? (nfull == 0 ? froff : 0) + (T - r_off[nfull]) : 0;
				const uint32_t fb = (uint32_t)((uint64_t)T * wid / NW), fe = (uint32_t)((uint64_t)T * (wid + 1) / NW);
				if (fb < fe) {
					uint32_t k = find_run(fb);
					for (uint32_t f0 = fb; f0 < fe; f0 += STEP) {
						tup_t t[RU];
						load_step(f0, fe, k, t);
#pragma unroll
						for (uint32_t u = 0; u < RU; u++) {
							const uint32_t f = f0 + u * HJ_WAVE + lane;
							if (f >= fe) continue;
							const uint64_t key = key_of(t[u]);
							if (mode == 1) tab.count(key, hs);
							else if (mode == 2) tab.place(key, pay_of(t[u]), hs);
							else tab.put(m + f, key, pay_of(t[u]), hs);
						}
					}
				}
				m += T;
				frc += nfull;
				froff = roff_new;
				__syncthreads();
				if (m >= cap || frc >= pr.r1) break;
			}
		};
		while (rc < pr.r1) {
			// ---- fill the table with up to `cap` build tuples
			tab.reset(cap, NT);
			__syncthreads();
			uint32_t m = 0;
			if (TWO) {
				uint64_t rc0 = rc;
				uint32_t roff0 = roff, m0 = 0;
				fill(1, rc0, roff0, m0);
				tab.starts(sred, NT);
				__syncthreads();
			}
			fill(TWO ? 2 : 0, rc, roff, m);
			if (LP) {
				if (rc < pr.r1) {
					oversize = true;  // the host recomputes this join exactly
					break;
				}
				tab.lp_insert(m, hs, NT, dup);
				__syncthreads();
			}
			// ---- probe the tile's runs, NT chunks at a time
			for (uint64_t kb = sb; m && kb < se; kb += NT) {
				const uint32_t T = stage(kb, min(se, kb + NT), p2, 0);
				__syncthreads();
				const uint32_t fb = (uint32_t)((uint64_t)T * wid / NW), fe = (uint32_t)((uint64_t)T * (wid + 1) / NW);
				if (fb < fe) {
					uint32_t k = find_run(fb);
					tup_t ta[RU] = {}, tb[RU] = {};
					// loads never under a condition (indices clamp to fe - 1): a
					// conditional prefetch makes the compiler wait at the merge
					// for the loads it has just issued
					load_step(fb, fe, k, ta);
					for (uint32_t f0 = fb; f0 < fe; f0 += 2 * STEP) {
						load_step(f0 + STEP, fe, k, tb);
						probe_step(f0, fe, ta);
						load_step(f0 + 2 * STEP, fe, k, ta);
						probe_step(f0 + STEP, fe, tb);  // lanes past fe skip
					}
				}
				__syncthreads();
			}
		}
		__syncthreads();
	}
	acc.finish();
	const unsigned fl = (__any(dup) ? HJ_DF_DUPKEY : 0u) | (__any(oversize) ? HJ_DF_OVERSIZE : 0u);
	if (fl && (threadIdx.x & (HJ_WAVE - 1)) == 0) atomicOr(&st->flags, fl);
	hj_block_add2(acc.cnt, acc.sum, red, &st->matched, &st->checksum);
}

// Narrow bucket-chain join: k_join2's work loop for the common case of a bc
// join over the 8-byte {u32 key, u32 row_id} tuples pass B writes when every
// key and row id fits 32 bits.  Its table keeps one 8-byte word per build
// tuple (key | payload << 32) with a u16 chain link and u32 bucket heads,
// 10 B per tuple + 4 B per bucket instead of 18 + 4, so three workgroups
// share a CU (more waves to cover the loads' latency), and it compares 32-bit
// keys.  A staged run is one {source, flat offset} word, and each lane keeps
// its current run's source and offset and the next run's offset in registers,
// so it reads LDS for addressing only when its next tuple lies in another run.
// Same table semantics as HJ_TK_BC (head insertion, every pair counted,
// bucket (key >> B) & (nb - 1), fills of at most `cap` build tuples joined
// one after the other), so the result is bc's.
// LDS: kv u64[cap] | heads u32[np2(cap)] | next u16[cap]
//
// LPT: the same loop with HJ_TK_LP's table (k_join2's lp): slots u32[np2(cap)]
// hold 1-based tuple indices, claimed by CAS in the fill's order after the
// whole partition is in LDS, linear probing from (key >> B) & (nb - 1), the
// first match only.  A second build tuple with the same key (HJ_DF_DUPKEY) or
// a partition larger than one fill (HJ_DF_OVERSIZE) raises a flag and the
// join is recomputed by the exact path, as after k_join2.
// LDS: kv u64[cap] | slots u32[np2(cap)]
template <int NT, bool LPT = false>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(6))) void k_join2n(const uint2 *__restrict__ P, const uint64_t *__restrict__ cstart,
                                               const uint16_t *__restrict__ run_t, uint32_t stride, uint32_t b1,
                                               const hj_pair *__restrict__ pairs,
                                               const uint64_t *__restrict__ tprefix, uint32_t npairs, uint32_t B,
                                               uint32_t cpt, uint32_t cap, hj_stats *st, bool narrow_ok) {
	if (!(narrow_ok && !(st->flags & HJ_DF_WIDE))) return;  // k_join2<1, ...> joins 16-byte tuples
	const uint32_t hs = B - b1;  // keys arrive shifted right by b1 bits (k_sortchunk_g)
	constexpr uint32_t NW = NT / HJ_WAVE;
	constexpr uint32_t RU = 4;               // tuples per lane per step
	constexpr uint32_t STEP = HJ_WAVE * RU;  // tuples per wave per step
	extern __shared__ uint64_t lds_u64[];
	__shared__ unsigned long long red[2 * NW];
	__shared__ uint32_t sred[NW + 1];
	__shared__ uint64_t s_item[4];
	__shared__ uint2 ent[NT + 1];  // staged run k: {first tuple, flat offset}; ent[NT].y = total
	__shared__ uint32_t pend[NT];  // staged probe run k: its end (tuple index)
	const uint32_t nb = hj_np2u(cap), hmask = nb - 1;
	uint64_t *kv = lds_u64;
	uint32_t *heads = (uint32_t *)(kv + cap);  // LPT: the slots
	uint16_t *next = (uint16_t *)(heads + nb);  // (not used with LPT)
	[[maybe_unused]] bool dup = false, oversize = false;
	const uint32_t lane = threadIdx.x & (HJ_WAVE - 1), wid = threadIdx.x / HJ_WAVE;
	const uint64_t total = tprefix[npairs];
	const uint32_t b2 = B - b1, m2 = (1u << b2) - 1;
	uint32_t cur = blockIdx.x & 7;  // this workgroup's queue (its XCD), as k_join2
	unsigned long long cnt = 0, sum = 0;

	// runs of digit p2 of chunks [c0, c1), c1 - c0 <= NT; the first run starts
	// `skip` tuples in.  Positions fit 32 bits: narrow tuples are written only
	// when |R| and |S| are below 2^31.
	auto stage = [&](uint64_t c0, uint64_t c1, uint32_t p2, uint32_t skip) -> uint32_t {
		const uint64_t c = c0 + threadIdx.x;
		uint32_t len = 0, src = 0;
		if (c < c1) {
			uint32_t a = run_t[(uint64_t)p2 * stride + c];
			len = run_t[(uint64_t)(p2 + 1) * stride + c] - a;
			if (threadIdx.x == 0) {
				a += skip;
				len -= skip;
			}
			src = (uint32_t)(cstart[c] + a);
		}
		uint32_t tot;
		const uint32_t ex = hj_block_exscan<uint32_t>(len, sred, &tot);
		ent[threadIdx.x] = make_uint2(src, ex);
		if (threadIdx.x == NT - 1) ent[NT] = make_uint2(0u, tot);
		return tot;
	};
	// a lane's position in the staged runs: run k holds flat indices
	// [off, noff) starting at tuple src
	struct cursor_t {
		uint32_t k, src, off, noff;
	};
	auto seek = [&](uint32_t f) -> cursor_t {
		uint32_t lo = 0, hi = NT;
		while (hi - lo > 1) {
			const uint32_t mid = (lo + hi) >> 1;
			if (ent[mid].y <= f) lo = mid; else hi = mid;
		}
		const uint2 e = ent[lo];
		return cursor_t{lo, e.x, e.y, ent[lo + 1].y};
	};
	// Load flat indices f0 + u*64 + lane (clamped to fe - 1: every lane issues
	// RU loads, none under a lane condition); the cursor only moves forward.
	auto load_step = [&](uint32_t f0, uint32_t fe, cursor_t &cs, uint2 (&t)[RU]) {
#pragma unroll
		for (uint32_t u = 0; u < RU; u++) {
			const uint32_t f = min(f0 + u * HJ_WAVE + lane, fe - 1);
			while (f >= cs.noff) {
				cs.k++;
				cs.off = cs.noff;
				cs.src = ent[cs.k].x;
				cs.noff = ent[cs.k + 1].y;
			}
			t[u] = hj_load8(P + cs.src + (f - cs.off));
		}
	};
	// The probe side is read in aligned pairs: the 16 bytes at tuple index
	// 2j hold tuples 2j and 2j+1, and a staged run [s, e) covers pairs s/2 to
	// (e-1)/2 (half the load instructions of 8-byte loads, none crossing a
	// 16-byte boundary); a pair's half outside the run (the lower one of the
	// first pair when s is odd, the upper one of the last when e is odd)
	// belongs to another run and is masked.  A staged run keeps its start s
	// in ent[k].x and its end e in pend[k].
	constexpr uint32_t RP = RU / 2;             // pairs per lane per step
	constexpr uint32_t PSTEP = HJ_WAVE * RP;    // pairs per wave per step
	auto stage_pairs = [&](uint64_t c0, uint64_t c1, uint32_t p2) -> uint32_t {
		const uint64_t c = c0 + threadIdx.x;
		uint32_t np = 0, s0 = 0, e = 0;
		if (c < c1) {
			const uint32_t a = run_t[(uint64_t)p2 * stride + c];
			const uint32_t len = run_t[(uint64_t)(p2 + 1) * stride + c] - a;
			s0 = (uint32_t)(cstart[c] + a);
			e = s0 + len;
			np = len ? ((e - 1) >> 1) - (s0 >> 1) + 1 : 0u;
		}
		uint32_t tot;
		const uint32_t ex = hj_block_exscan<uint32_t>(np, sred, &tot);
		ent[threadIdx.x] = make_uint2(s0, ex);
		pend[threadIdx.x] = e;
		if (threadIdx.x == NT - 1) ent[NT] = make_uint2(0u, tot);
		return tot;
	};
	// a lane's pair cursor: run k (tuples [s, e)) holds flat pairs
	// [off, noff), the first of them the pair at tuple index 2*(s/2)
	struct pcursor_t {
		uint32_t k, s, off, noff, e;
	};
	auto pseek = [&](uint32_t f) -> pcursor_t {
		uint32_t lo = 0, hi = NT;
		while (hi - lo > 1) {
			const uint32_t mid = (lo + hi) >> 1;
			if (ent[mid].y <= f) lo = mid; else hi = mid;
		}
		const uint2 x = ent[lo];
		return pcursor_t{lo, x.x, x.y, ent[lo + 1].y, pend[lo]};
	};
	// pairs f0 + u*64 + lane, clamped to fe - 1 (every lane loads); bits 2u
	// and 2u+1 of ok: the pair's lower and upper tuple lie in its run
	auto load_pairs = [&](uint32_t f0, uint32_t fe, pcursor_t &cs, uint4 (&t)[RP], uint32_t &ok) {
		ok = 0;
#pragma unroll
		for (uint32_t u = 0; u < RP; u++) {
			const uint32_t f = min(f0 + u * HJ_WAVE + lane, fe - 1);
			while (f >= cs.noff) {
				cs.k++;
				cs.off = cs.noff;
				cs.s = ent[cs.k].x;
				cs.noff = ent[cs.k + 1].y;
				cs.e = pend[cs.k];
			}
			const uint32_t t2 = 2 * ((cs.s >> 1) + (f - cs.off));  // tuple index of the pair's lower half
			t[u] = hj_load((const uint4 *)(P + t2));
			ok |= ((uint32_t)(t2 >= cs.s) | ((uint32_t)(t2 + 1 < cs.e) << 1)) << (2 * u);
		}
	};
	auto probe_pairs = [&](uint32_t f0, uint32_t fe, const uint4 (&t)[RP], uint32_t ok) {
		uint32_t c[2 * RP], key[2 * RP], rid[2 * RP];
#pragma unroll
		for (uint32_t u = 0; u < RP; u++) {
			const bool live = f0 + u * HJ_WAVE + lane < fe;
			key[2 * u] = t[u].x;
			rid[2 * u] = t[u].y;
			key[2 * u + 1] = t[u].z;
			rid[2 * u + 1] = t[u].w;
			// bc: the chain's first entry; lp: the home slot's index + 1
			const uint32_t h0 = (t[u].x >> hs) & hmask, h1 = (t[u].z >> hs) & hmask;
			c[2 * u] = live && ((ok >> (2 * u)) & 1u) ? (LPT ? h0 + 1 : heads[h0]) : 0u;
			c[2 * u + 1] = live && ((ok >> (2 * u + 1)) & 1u) ? (LPT ? h1 + 1 : heads[h1]) : 0u;
		}
		if constexpr (LPT) {
			// c[u] = the next slot's index + 1 (0: done); at most nb slots
			// (the table always has an empty one: m <= cap <= nb)
			bool any = true;
			for (uint32_t step = 0; any && step < nb; step++) {
				any = false;
#pragma unroll
				for (uint32_t u = 0; u < 2 * RP; u++) {
					if (!c[u]) continue;
					const uint32_t o = heads[c[u] - 1];
					if (!o) {
						c[u] = 0;
						continue;
					}
					const uint64_t e = kv[o - 1];
					if ((uint32_t)e == key[u]) {
						cnt++;
						sum += (e >> 32) + rid[u];
						c[u] = 0;
						continue;
					}
					c[u] = (c[u] & hmask) + 1;  // slot c[u] - 1 + 1
					any = true;
				}
			}
			return;
		}
		bool any = true;
		while (any) {
			any = false;
#pragma unroll
			for (uint32_t u = 0; u < 2 * RP; u++) {
				if (!c[u]) continue;
				const uint64_t e = kv[c[u] - 1];
				if ((uint32_t)e == key[u]) {
					cnt++;
					sum += (e >> 32) + rid[u];
				}
				c[u] = next[c[u] - 1];
				any |= c[u] != 0;
			}
		}
	};

	for (;;) {
		if (threadIdx.x == 0) {
			uint64_t w, tix;
			uint32_t lo;
			hj_claim_item(tprefix, npairs, total, st->xq, cur, w, lo, tix);
			s_item[0] = w;
			s_item[1] = lo;
			s_item[2] = tix;
		}
		__syncthreads();
		const uint64_t w = s_item[0];
		const uint32_t q = (uint32_t)s_item[1];
		const uint64_t tix = s_item[2];
		__syncthreads();
		if (w >= total) break;
		const hj_pair pr = pairs[q];
		const uint32_t p2 = q & m2;
		const uint64_t sb = pr.s0 + tix * cpt, se = min(pr.s1, sb + cpt);
		uint64_t rc = pr.r0;
		uint32_t roff = 0;  // tuples of run rc already taken by an earlier fill
		while (rc < pr.r1) {
			// ---- fill the table with up to `cap` build tuples
			for (uint32_t i = threadIdx.x; i < nb; i += NT) heads[i] = 0;
			__syncthreads();
			uint32_t m = 0;
			for (;;) {
				const uint64_t c1 = min(pr.r1, rc + NT);
				const uint32_t n = (uint32_t)(c1 - rc);
				const uint32_t tot = stage(rc, c1, p2, roff);
				__syncthreads();
				const uint32_t T = min(tot, cap - m);
				// runs taken whole form a prefix; the next one may be taken in part
				const uint32_t nfull = __syncthreads_count(threadIdx.x < n && ent[threadIdx.x + 1].y <= T);
				const uint32_t roff_new = nfull < n ? (nfull == 0 ? roff : 0) + (T - ent[nfull].y) : 0;
				const uint32_t fb = (uint32_t)((uint64_t)T * wid / NW), fe = (uint32_t)((uint64_t)T * (wid + 1) / NW);
				if (fb < fe) {
					cursor_t cs = seek(fb);
					for (uint32_t f0 = fb; f0 < fe; f0 += STEP) {
						uint2 t[RU];
						load_step(f0, fe, cs, t);
#pragma unroll
						for (uint32_t u = 0; u < RU; u++) {
							const uint32_t f = f0 + u * HJ_WAVE + lane;
							if (f >= fe) continue;
							const uint32_t i = m + f;
							kv[i] = (uint64_t)t[u].x | (uint64_t)t[u].y << 32;
							if constexpr (!LPT) next[i] = (uint16_t)atomicExch(&heads[(t[u].x >> hs) & hmask], i + 1);
						}
					}
				}
				m += T;
				rc += nfull;
				roff = roff_new;
				__syncthreads();
				if (m >= cap || rc >= pr.r1) break;
			}
			if constexpr (LPT) {
				if (rc < pr.r1) {
					oversize = true;  // the host recomputes this join exactly
					break;
				}
				// claim slots for the m tuples in LDS (every thread saw the
				// fill's last barrier, so kv[0..m) is complete)
				for (uint32_t i = threadIdx.x; i < m; i += NT) {
					const uint32_t key = (uint32_t)kv[i];
					uint32_t p = (key >> hs) & hmask;
					for (uint32_t step = 0; step < nb; step++) {
						const uint32_t old = atomicCAS(&heads[p], 0u, i + 1);
						if (old == 0) break;
						if ((uint32_t)kv[old - 1] == key) {
							dup = true;
							break;
						}
						p = (p + 1) & hmask;
					}
				}
				__syncthreads();
			}
			// ---- probe the tile's runs, NT chunks at a time, in pairs
			for (uint64_t kb = sb; m && kb < se; kb += NT) {
				const uint32_t T = stage_pairs(kb, min(se, kb + NT), p2);
				__syncthreads();
				const uint32_t fb = (uint32_t)((uint64_t)T * wid / NW), fe = (uint32_t)((uint64_t)T * (wid + 1) / NW);
				if (fb < fe) {
					pcursor_t cs = pseek(fb);
					uint4 ta[RP], tb[RP];
					uint32_t oa, ob;
					// unconditional prefetch (indices clamp to fe - 1), see k_join2
					load_pairs(fb, fe, cs, ta, oa);
					for (uint32_t f0 = fb; f0 < fe; f0 += 2 * PSTEP) {
						load_pairs(f0 + PSTEP, fe, cs, tb, ob);
						probe_pairs(f0, fe, ta, oa);
						load_pairs(f0 + 2 * PSTEP, fe, cs, ta, oa);
						probe_pairs(f0 + PSTEP, fe, tb, ob);  // lanes past fe skip
					}
				}
				__syncthreads();
			}
		}
		__syncthreads();
	}
	if constexpr (LPT) {
		const unsigned fl = (__any(dup) ? HJ_DF_DUPKEY : 0u) | (__any(oversize) ? HJ_DF_OVERSIZE : 0u);
		if (fl && (threadIdx.x & (HJ_WAVE - 1)) == 0) atomicOr(&st->flags, fl);
	}
	hj_block_add2(cnt, sum, red, &st->matched, &st->checksum);
}

// LDS bytes of k_join2n's table for `cap` build tuples (lp: no chain links)
static size_t join2n_lds(uint32_t cap, bool lp = false) {
	return (size_t)cap * (lp ? 8 : 10) + (size_t)hj_np2u(cap) * 4 + 8;
}

// ------------------------------------------------------------ host side
struct part_cfg {
	int nt, items, tile;
};

// Tile shapes (threads x tuples per thread).  16-byte tuples have several
// shapes for tuning sweeps; wider tuples use one 32 KB tile.
static const part_cfg k_shapes16[] = {{512, 4, 2048}, {1024, 4, 4096}, {512, 8, 4096},
                                      {256, 8, 2048}, {1024, 2, 2048}, {256, 16, 4096}};

static part_cfg cfg_for(uint32_t wv, uint32_t variant = 0) {
	variant &= 15;  // shape field
	if (wv == 1) return k_shapes16[variant < 6 ? variant : 0];
	const int nt = wv >= 8 ? 256 : 512;
	int items = (int)(32768 / (16 * wv) / nt);
	if (items < 1) items = 1;
	return part_cfg{nt, items, nt * items};
}

// LDS of k_scatter staging tb-byte tuples (16 * wv, or 8 for narrow output)
static size_t scatter_lds_b(const part_cfg &c, uint32_t tb, uint32_t F, uint32_t lt) {
	const size_t xt = (size_t)c.tile + (size_t)F * (lt - 1);
	return xt * tb + (size_t)F * (lt - 1) * tb + (size_t)F * 20 + ((xt + 3) & ~(size_t)1) * 2 + 64 * 4;
}
static size_t scatter_lds(const part_cfg &c, uint32_t wv, uint32_t F, uint32_t lt) {
	return scatter_lds_b(c, 16 * wv, F, lt);
}

// Write-combining line of the scatter (tuples of tb bytes): the largest of
// 128 / tb, 64 / tb (at least one tuple) whose carries still let two
// workgroups share a CU; 1 when neither does.
static uint32_t scatter_lt_b(const part_cfg &c, uint32_t tb, uint32_t F, size_t lds_cu) {
	for (uint32_t line = 128; line >= 64; line >>= 1) {
		const uint32_t lt = line / tb;
		if (lt >= 2 && scatter_lds_b(c, tb, F, lt) + 1024 <= lds_cu / 2) return lt;
	}
	return 1;
}
static uint32_t scatter_lt(const part_cfg &c, uint32_t wv, uint32_t F, size_t lds_cu) {
	return scatter_lt_b(c, 16 * wv, F, lds_cu);
}

// Static LDS of a kernel (its __shared__ arrays), so that static + dynamic
// never exceeds the device's per-workgroup LDS.
static size_t static_lds(const void *kfn) {
	hipFuncAttributes a{};
	if (hipFuncGetAttributes(&a, kfn) != hipSuccess) {
		(void)hipGetLastError();
		return 16384;  // conservative if the query fails
	}
	return a.sharedSizeBytes;
}

// Allow a kernel the device's whole LDS beyond its static arrays.  Every
// caller sets the same value, so launches from several host threads (server
// lanes, batch slots) cannot lower the limit under one another.  Attributes
// are not stream work: callers skip this while capturing a graph (a batch
// sets them before it starts capturing).
static void allow_lds(hj_ctx_t *ctx, const void *kfn) {
	const size_t dyn = ctx->prop.sharedMemPerBlock - std::min<size_t>(ctx->prop.sharedMemPerBlock, static_lds(kfn));
	if (hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn) != hipSuccess)
		(void)hipGetLastError();
}

// Workgroups of `kfn` that are resident at once on the device: the grid of a
// persistent kernel, so that no workgroup waits for another to finish before
// it starts its share.
static unsigned resident_grid(hj_ctx_t *ctx, const void *kfn, int nt, size_t lds) {
	int nb = 0;
	if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kfn, nt, lds) != hipSuccess || nb < 1) {
		(void)hipGetLastError();
		nb = 1;
	}
	return (unsigned)nb * (unsigned)ctx->num_cu;
}

static bool capturing(hj_ctx_t *ctx) {
	hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
	if (hipStreamIsCapturing(ctx->stream, &cs) != hipSuccess) (void)hipGetLastError();
	return cs != hipStreamCaptureStatusNone;
}

// Launchers take the stats block explicitly: a context's single joins use
// ctx->stats, a batch uses its own.  The kernels read the chunk count there.
template <int WV>
static void launch_hist(hj_ctx_t *ctx, unsigned grid, const hj_seg *segs, const uint32_t *cmap, const hj_stats *st,
                        uint64_t ch, uint32_t shift, uint32_t bits, uint64_t *hist, bool detect) {
	if constexpr (WV == 1) {
		if (detect) {
			hipLaunchKernelGGL((k_hist<WV, true>), dim3(grid), dim3(512), (size_t)4 << bits, ctx->stream, segs, cmap,
			                   (hj_stats *)st, ch, shift, bits, hist);
			return;
		}
	}
	hipLaunchKernelGGL((k_hist<WV, false>), dim3(grid), dim3(512), (size_t)4 << bits, ctx->stream, segs, cmap,
	                   (hj_stats *)st, ch, shift, bits, hist);
}

template <int WV, int NT, int ITEMS>
static void launch_scatter_t(hj_ctx_t *ctx, unsigned grid, size_t lds, const hj_seg *segs, const uint32_t *cmap,
                             const hj_stats *st, uint64_t ch, uint32_t shift, uint32_t bits, uint32_t lt,
                             const uint64_t *scan, void *out, bool narrow_ok = false) {
	auto kfn = k_scatter<WV, NT, ITEMS, false>;
	if (!capturing(ctx)) allow_lds(ctx, (const void *)kfn);
	hipLaunchKernelGGL(kfn, dim3(grid), dim3(NT), lds, ctx->stream, segs, cmap, st, ch, shift, bits, lt, scan,
	                   (uint4 *)out, narrow_ok);
}

// Write-combining line of a partition pass: scatter_lt's choice, or the one
// part_variant bits 4-6 force for tuning sweeps.
static uint32_t pass_line(hj_ctx_t *ctx, const part_cfg &c, uint32_t wv, uint32_t F, uint32_t variant) {
	if (variant & 0x70) return 1u << (((variant >> 4) & 7) - 1);
	return scatter_lt(c, wv, F, ctx->prop.sharedMemPerBlock);
}

// Narrow output of a batch pass (16-byte tuples): k_scatter<1, 512, 4, true>
// with 8-byte staging; its line and LDS.
static constexpr part_cfg k_narrow_scatter{512, 4, 2048};

// narrow: the pass may write 8-byte tuples (16-byte input; the histogram
// pass decides on the device, st->flags HJ_DF_WIDE) for a consumer that
// reads both layouts (the batch join); both scatter instances are launched.
static int launch_pass(hj_ctx_t *ctx, uint32_t wv, uint32_t variant, unsigned grid, const hj_seg *segs,
                       const uint32_t *cmap, uint64_t ch, uint32_t shift, uint32_t bits, uint64_t *hist,
                       uint64_t hist_len, void *out, const hj_stats *st, uint64_t *scan_tmp, bool narrow = false) {
	const part_cfg c = cfg_for(wv, variant);
	const uint32_t F = 1u << bits;
	const uint32_t lt = pass_line(ctx, c, wv, F, variant);
	const size_t lds = scatter_lds(c, wv, F, lt);
	if (lds + 1024 > ctx->prop.sharedMemPerBlock)
		return hj_fail(ctx, HJ_EINVAL, "%u radix bits in one pass need %zu B of LDS (device: %zu)", bits, lds,
		               (size_t)ctx->prop.sharedMemPerBlock);
	narrow = narrow && wv == 1;
	if (grid == 0) return HJ_OK;
	hj_kt_begin(ctx, HJ_K_HIST);
	switch (wv) {
	case 1: launch_hist<1>(ctx, grid, segs, cmap, st, ch, shift, bits, hist, narrow); break;
	case 2: launch_hist<2>(ctx, grid, segs, cmap, st, ch, shift, bits, hist, false); break;
	case 4: launch_hist<4>(ctx, grid, segs, cmap, st, ch, shift, bits, hist, false); break;
	case 8: launch_hist<8>(ctx, grid, segs, cmap, st, ch, shift, bits, hist, false); break;
	default: launch_hist<16>(ctx, grid, segs, cmap, st, ch, shift, bits, hist, false); break;
	}
	hj_kt_end(ctx);
	hj_kt_begin(ctx, HJ_K_SCAN);
	int rc = scan_u64(ctx, hist, hist_len, scan_tmp);
	hj_kt_end(ctx);
	if (rc) return rc;
	hj_kt_begin(ctx, HJ_K_SCATTER);
	if (narrow) {
		// 8-byte instance (returns unless every tuple is narrow); tiles of
		// 2048 tuples, which every chunk size here is a multiple of
		const uint32_t ltn = (variant & 0x70) ? lt : scatter_lt_b(k_narrow_scatter, 8, F, ctx->prop.sharedMemPerBlock);
		const size_t ldsn = scatter_lds_b(k_narrow_scatter, 8, F, ltn);
		auto kn = k_scatter<1, 512, 4, true>;
		if (!capturing(ctx)) allow_lds(ctx, (const void *)kn);
		hipLaunchKernelGGL(kn, dim3(grid), dim3(512), ldsn, ctx->stream, segs, cmap, st, ch, shift, bits, ltn, hist,
		                   (uint4 *)out, true);
	}
	switch (wv) {
	case 1:
		switch ((variant & 15) < 6 ? (variant & 15) : 0) {
#define S1(N, I) launch_scatter_t<1, N, I>(ctx, grid, lds, segs, cmap, st, ch, shift, bits, lt, hist, out, narrow)
		case 2: S1(512, 8); break;
		case 3: S1(256, 8); break;
		case 4: S1(1024, 2); break;
		case 5: S1(256, 16); break;
		case 1: S1(1024, 4); break;
		default: S1(512, 4); break;
#undef S1
		}
		break;
	case 2: launch_scatter_t<2, 512, 2>(ctx, grid, lds, segs, cmap, st, ch, shift, bits, lt, hist, out); break;
	case 4: launch_scatter_t<4, 512, 1>(ctx, grid, lds, segs, cmap, st, ch, shift, bits, lt, hist, out); break;
	case 8: launch_scatter_t<8, 256, 1>(ctx, grid, lds, segs, cmap, st, ch, shift, bits, lt, hist, out); break;
	default: launch_scatter_t<16, 256, 1>(ctx, grid, lds, segs, cmap, st, ch, shift, bits, lt, hist, out); break;
	}
	hj_kt_end(ctx);
	HJ_HIP(ctx, hipGetLastError());
	return HJ_OK;
}

// instances of one kernel template for every tuple width and table kind
#define HJ_TK_SWITCH(tk, X)              \
	switch (tk) {                        \
	case HJ_TK_LP: X(HJ_TK_LP); break;   \
	case HJ_TK_SC: X(HJ_TK_SC); break;   \
	case HJ_TK_HR: X(HJ_TK_HR); break;   \
	case HJ_TK_HRO: X(HJ_TK_HRO); break; \
	default: X(HJ_TK_BC); break;         \
	}

static size_t join_static_lds(int tk, bool two_pass) {
	// every width of one kernel template has the same __shared__ arrays
	size_t r = 16384;
#define S(K) \
	r = two_pass ? static_lds((const void *)k_join2<1, 512, K, false, false>) \
	             : static_lds((const void *)k_join<1, 256, K, false>)
	HJ_TK_SWITCH(tk, S)
#undef S
	return r;
}

// Dynamic-LDS attribute of the default-shape scatter kernel for width wv.
static void set_scatter_attr(hj_ctx_t *ctx, uint32_t wv) {
	const void *k = nullptr;
	switch (wv) {
	case 1:
		// every tile shape a variant may pick, and the narrow instance
		allow_lds(ctx, (const void *)k_scatter<1, 1024, 4, false>);
		allow_lds(ctx, (const void *)k_scatter<1, 512, 8, false>);
		allow_lds(ctx, (const void *)k_scatter<1, 256, 8, false>);
		allow_lds(ctx, (const void *)k_scatter<1, 1024, 2, false>);
		allow_lds(ctx, (const void *)k_scatter<1, 256, 16, false>);
		allow_lds(ctx, (const void *)k_scatter<1, 512, 4, true>);
		k = (const void *)k_scatter<1, 512, 4, false>;
		break;
	case 2: k = (const void *)k_scatter<2, 512, 2, false>; break;
	case 4: k = (const void *)k_scatter<4, 512, 1, false>; break;
	case 8: k = (const void *)k_scatter<8, 256, 1, false>; break;
	default: k = (const void *)k_scatter<16, 256, 1, false>; break;
	}
	allow_lds(ctx, k);
}

// LDS bytes of a bc or lp join table holding `cap` build tuples (batches).
static size_t join_lds(uint32_t cap, bool lp) { return hj_tab_lds(lp ? HJ_TK_LP : HJ_TK_BC, cap, 1); }

// Joins that emit row-id pairs (late materialisation) run on 16-byte
// key/row-id tuples only; every other instance counts and sums.
template <int WV, int TK>
static void launch_join_t(hj_ctx_t *ctx, unsigned grid, size_t lds, const void *Rb, const void *Sb, const hj_pair *pairs,
                          const uint64_t *tpre, uint32_t npairs, uint32_t B, uint64_t s_tile, uint32_t cap) {
	auto kfn = k_join<WV, 256, TK, false>;
	if constexpr (WV == 1 && TK != HJ_TK_HRO)
		if (ctx->emit_pairs) kfn = k_join<1, 256, TK, true>;
	allow_lds(ctx, (const void *)kfn);
	hipLaunchKernelGGL(kfn, dim3(grid), dim3(256), lds, ctx->stream, (const uint4 *)Rb, (const uint4 *)Sb, pairs,
	                   tpre, npairs, B, s_tile, cap, (hj_stats *)ctx->stats.p, ctx->emit_pairs,
	                   (unsigned long long)ctx->emit_cap, ctx->emit_chunk);
}

// hro joins blocks of tuples of at most 64 bytes (phj.c:880), so it has no
// instances for wider tuples
template <int W, int TK>
struct hj_tk_ok {
	static constexpr bool value = TK != HJ_TK_HRO || W <= 4;
};

static void launch_join(hj_ctx_t *ctx, uint32_t wv, int tk, unsigned grid, size_t lds, const void *Rb,
                        const void *Sb, const hj_pair *pairs, const uint64_t *tpre, uint32_t npairs, uint32_t B,
                        uint64_t s_tile, uint32_t cap) {
#define J(W, K) \
	if constexpr (hj_tk_ok<W, K>::value) launch_join_t<W, K>(ctx, grid, lds, Rb, Sb, pairs, tpre, npairs, B, s_tile, cap)
#define JW(K)                    \
	switch (wv) {                \
	case 1: J(1, K); break;      \
	case 2: J(2, K); break;      \
	case 4: J(4, K); break;      \
	case 8: J(8, K); break;      \
	case 16: J(16, K); break;    \
	default: J(32, K); break;    \
	}
	HJ_TK_SWITCH(tk, JW)
#undef JW
#undef J
}

template <int WV, int NT, int ITEMS>
static void launch_sortchunk_t(hj_ctx_t *ctx, unsigned grid, size_t lds, const hj_seg *segs, const uint32_t *cmap,
                               uint32_t shift, uint32_t bits, void *out, uint16_t *run, uint64_t *cst) {
	auto kfn = k_sortchunk<WV, NT, ITEMS>;
	allow_lds(ctx, (const void *)kfn);
	grid = std::max(1u, std::min(grid, resident_grid(ctx, (const void *)kfn, NT, lds)));
	hipLaunchKernelGGL(kfn, dim3(grid), dim3(NT), lds, ctx->stream, segs, cmap, (hj_stats *)ctx->stats.p, shift,
	                   bits, (uint4 *)out, run, cst, ctx->narrow_ok, ctx->dense_ok, ctx->tile_narrow);
}

// Pass-A tile of 16-byte tuples (threads x tuples per thread): long enough
// that a run of one first digit in a tile is 16-32 tuples, since pass B
// gathers those runs (2048 tuples up to 7 first-pass bits; above, 8192 by
// 1024 threads x 8 for dense tuples up to 11 bits and 4096 by 1024 x 4 at
// 12, 4096 by 512 threads with 8 loads each for the heads of wider tuples
// read in place, whose loads are spread further); part_variant bits 12-13
// force 1: 512x4, 2: 1024x4, 3: 512x8, bit 16 1024x8.  Wider tuples use
// cfg_for(wv, 0).
static part_cfg cfg_passA(uint32_t wv, uint32_t variant, uint32_t b1, uint32_t in_wv) {
	if (wv != 1) return cfg_for(wv, 0);
	if ((variant >> 16) & 1) return part_cfg{1024, 8, 8192};  // part_variant bit 16
	switch ((variant >> 12) & 3) {
	case 1: return part_cfg{512, 4, 2048};
	case 2: return part_cfg{1024, 4, 4096};
	case 3: return part_cfg{512, 8, 4096};
	default:
		if (b1 < 8) return part_cfg{512, 4, 2048};
		if (in_wv > 1) return part_cfg{512, 8, 4096};
		// 8192 tuples while the tile and 2^b1 counters fit the LDS: pass B's
		// runs get twice as long (C2 2.81 -> 2.71 ms)
		return b1 <= 11 ? part_cfg{1024, 8, 8192} : part_cfg{1024, 4, 4096};
	}
}

static void launch_sortchunk(hj_ctx_t *ctx, uint32_t wv, const part_cfg &pc, unsigned grid, size_t lds,
                             const hj_seg *segs, const uint32_t *cmap, uint32_t shift, uint32_t bits, void *out,
                             uint16_t *run, uint64_t *cst) {
	// shapes must match cfg_passA
	switch (wv) {
	case 1:
		if (pc.nt == 1024 && pc.items == 8)
			launch_sortchunk_t<1, 1024, 8>(ctx, grid, lds, segs, cmap, shift, bits, out, run, cst);
		else if (pc.nt == 1024)
			launch_sortchunk_t<1, 1024, 4>(ctx, grid, lds, segs, cmap, shift, bits, out, run, cst);
		else if (pc.items == 8)
			launch_sortchunk_t<1, 512, 8>(ctx, grid, lds, segs, cmap, shift, bits, out, run, cst);
		else
			launch_sortchunk_t<1, 512, 4>(ctx, grid, lds, segs, cmap, shift, bits, out, run, cst);
		break;
	case 2: launch_sortchunk_t<2, 512, 2>(ctx, grid, lds, segs, cmap, shift, bits, out, run, cst); break;
	case 4: launch_sortchunk_t<4, 512, 1>(ctx, grid, lds, segs, cmap, shift, bits, out, run, cst); break;
	case 8: launch_sortchunk_t<8, 256, 1>(ctx, grid, lds, segs, cmap, shift, bits, out, run, cst); break;
	case 16: launch_sortchunk_t<16, 256, 1>(ctx, grid, lds, segs, cmap, shift, bits, out, run, cst); break;
	default: launch_sortchunk_t<32, 256, 1>(ctx, grid, lds, segs, cmap, shift, bits, out, run, cst); break;
	}
}

template <int WV, int NT, int ITEMS, bool DENSE, bool NIN>
static void launch_sortchunk_g_i(hj_ctx_t *ctx, unsigned grid, size_t lds, const hj_gtile *gt, const uint32_t *hist,
                                 const uint16_t *beg_t, uint32_t b1, uint32_t bits, const void *in, void *out,
                                 uint16_t *run_t, uint32_t rstride, uint64_t *cst, uint32_t tileA, uint32_t pmode) {
	// grid: all resident workgroups up to the cap `grid`, a multiple of the
	// 8 XCDs (the kernel's mapping); DENSE instances launch half as many
	// again, and use them only for mostly positional inputs
	const unsigned grid2 = std::max(8u, grid & ~7u);
	if (DENSE && grid != ~0u) grid += grid / 2;
	auto kfn = k_sortchunk_g<WV, NT, ITEMS, NIN, DENSE>;
	// the all-narrow instance's LDS tile holds 8-byte tuples
	const size_t l = NIN ? lds - (size_t)NT * ITEMS * 8 : lds;
	allow_lds(ctx, (const void *)kfn);
	grid = std::max(8u, std::min(grid, resident_grid(ctx, (const void *)kfn, NT, l)) & ~7u);
	hipLaunchKernelGGL(kfn, dim3(grid), dim3(NT), l, ctx->stream, gt, hist, beg_t, (const hj_stats *)ctx->stats.p, b1,
	                   bits, (const uint4 *)in, (uint4 *)out, run_t, rstride, cst, ctx->narrow_ok,
	                   (const uint8_t *)ctx->tile_narrow, tileA, std::min(grid2, grid), pmode);
}

// Which instance does the work is known on the device only: the all-narrow
// (NIN) one when no pass-A tile was wide, the 16-byte one otherwise; when
// pass A may have written positional words, the all-narrow work goes to the
// instance that decodes them if it did and to the plain one if not (which
// is faster on 8-byte sources).  The others return at once.
template <int WV, int NT, int ITEMS>
static void launch_sortchunk_g_t(hj_ctx_t *ctx, unsigned grid, size_t lds, const hj_gtile *gt, const uint32_t *hist,
                                 const uint16_t *beg_t, uint32_t b1, uint32_t bits, const void *in, void *out,
                                 uint16_t *run_t, uint32_t rstride, uint64_t *cst, uint32_t tileA) {
#define GI(D, N, M) \
	launch_sortchunk_g_i<WV, NT, ITEMS, D, N>(ctx, grid, lds, gt, hist, beg_t, b1, bits, in, out, run_t, rstride, cst, \
	                                          tileA, M)
	if constexpr (WV == 1) {
		if (ctx->dense_ok) {
			GI(true, true, 1);
			GI(false, true, 2);
			GI(true, false, 0);
			return;
		}
		if (ctx->narrow_ok) GI(false, true, 0);
	}
	GI(false, false, 0);
#undef GI
}

// LDS of k_sortchunk_g: tile, three source-offset buffers, counters
static size_t gathersort_lds(const part_cfg &c, uint32_t wv, uint32_t F) {
	return (size_t)c.tile * wv * 16 + (size_t)c.tile * 12 + (2 * (size_t)F + 1) * 4 + 64 * 4;
}

static void launch_sortchunk_g(hj_ctx_t *ctx, uint32_t wv, const part_cfg &pc, unsigned grid, size_t lds,
                               const hj_gtile *gt, const uint32_t *hist, const uint16_t *beg_t, uint32_t b1,
                               uint32_t bits, const void *in, void *out, uint16_t *run_t, uint32_t rstride, uint64_t *cst,
                               uint32_t tileA) {
	// shapes must match the pass-B tile (cfg_for(wv, 0), or 1024 x 4)
#define G(W, N, I) \
	launch_sortchunk_g_t<W, N, I>(ctx, grid, lds, gt, hist, beg_t, b1, bits, in, out, run_t, rstride, cst, tileA)
	switch (wv) {
	case 1:
		if (pc.nt == 1024) G(1, 1024, 4);
		else G(1, 512, 4);
		break;
	case 2: G(2, 512, 2); break;
	case 4: G(4, 512, 1); break;
	case 8: G(8, 256, 1); break;
	case 16: G(16, 256, 1); break;
	default: G(32, 256, 1); break;
	}
#undef G
}

template <int WV, int TK>
static void launch_join2_t(hj_ctx_t *ctx, unsigned grid, size_t lds, const void *P, const uint64_t *cst,
                           const uint16_t *run_t, uint32_t stride, uint32_t b1, const hj_pair *pairs,
                           const uint64_t *tpre, uint32_t npairs, uint32_t B, uint32_t cpt, uint32_t cap) {
	auto launch = [&](auto kfn) {
		allow_lds(ctx, (const void *)kfn);
		hipLaunchKernelGGL(kfn, dim3(grid), dim3(512), lds, ctx->stream, (const uint4 *)P, cst, run_t, stride, b1,
		                   pairs, tpre, npairs, B, cpt, cap, (hj_stats *)ctx->stats.p, ctx->emit_pairs,
		                   (unsigned long long)ctx->emit_cap, ctx->emit_chunk, ctx->narrow_ok);
	};
	if constexpr (WV == 1) {
		// pass B's layout is decided on the device: one of the two returns
		const bool emit = TK != HJ_TK_HRO && ctx->emit_pairs;
		if (emit) {
			if constexpr (TK != HJ_TK_HRO) {
				if (ctx->narrow_ok) launch(k_join2<1, 512, TK, true, true>);
				launch(k_join2<1, 512, TK, true, false>);
			}
		} else {
			if (ctx->narrow_ok) {
				if constexpr (TK == HJ_TK_BC || TK == HJ_TK_LP) {
					// the compact table: as many workgroups per CU as its LDS allows
					constexpr bool L = TK == HJ_TK_LP;
					auto kfn = k_join2n<512, L>;
					const size_t nl = join2n_lds(cap, L), all = nl + static_lds((const void *)kfn);
					const unsigned per_cu = (unsigned)std::max<size_t>(1, std::min<size_t>(4, (160 * 1024) / all));
					const unsigned ng = (unsigned)std::max<uint64_t>(
					    1, std::min<uint64_t>((uint64_t)ctx->num_cu * per_cu, (uint64_t)npairs * 64));
					allow_lds(ctx, (const void *)kfn);
					hipLaunchKernelGGL(kfn, dim3(ng), dim3(512), nl, ctx->stream, (const uint2 *)P, cst, run_t, stride, b1,
					                   pairs, tpre, npairs, B, cpt, cap, (hj_stats *)ctx->stats.p, ctx->narrow_ok);
				} else {
					launch(k_join2<1, 512, TK, false, true>);
				}
			}
			launch(k_join2<1, 512, TK, false, false>);
		}
	} else {
		launch(k_join2<WV, 512, TK, false, false>);
	}
}

static void launch_join2(hj_ctx_t *ctx, uint32_t wv, int tk, unsigned grid, size_t lds, const void *P,
                         const uint64_t *cst, const uint16_t *run_t, uint32_t stride, uint32_t b1,
                         const hj_pair *pairs, const uint64_t *tpre, uint32_t npairs, uint32_t B, uint32_t cpt,
                         uint32_t cap) {
#define J(W, K)                                                                                                    \
	if constexpr (hj_tk_ok<W, K>::value)                                                                           \
	launch_join2_t<W, K>(ctx, grid, lds, P, cst, run_t, stride, b1, pairs, tpre, npairs, B, cpt, cap)
#define JW(K)                    \
	switch (wv) {                \
	case 1: J(1, K); break;      \
	case 2: J(2, K); break;      \
	case 4: J(4, K); break;      \
	case 8: J(8, K); break;      \
	case 16: J(16, K); break;    \
	default: J(32, K); break;    \
	}
	HJ_TK_SWITCH(tk, JW)
#undef JW
#undef J
}

// Default radix configuration for |R| (the size-class autotuner's base
// rule): aim for ~2K build tuples per final partition so that the LDS table
// stays near 40 KB; two passes from 9 bits on.  Measured at |S| = 4..64 |R|:
// one pass wins up to 8 bits (|R| <= 2^19), two passes from 9 (2^20 x 2^24:
// 0.34 vs 0.52 ms; 2^22 x 2^28: 2.9 vs 11.0 ms), where a one-pass scatter's
// runs get short.
static uint32_t default_passes(uint32_t b) { return b == 0 ? 0 : (b > 8 ? 2 : 1); }
void hj_radix_default(uint64_t nr, uint32_t *bits, uint32_t *passes) {
	uint32_t b = 0;
	while (b < 18 && (nr >> b) > 2048) b++;
	*bits = b;
	*passes = default_passes(b);
}

void hj_radix_config(hj_ctx_t *ctx, uint64_t nr, uint32_t *bits, uint32_t *passes, uint32_t *pass1) {
	const uint32_t c = hj_log2_ceil(nr ? nr : 1);
	if (pass1) *pass1 = 0;
	if (ctx && ctx->tuned && c < 40 && ctx->tune_bits[c] != 0xff) {
		*bits = ctx->tune_bits[c];
		*passes = ctx->tune_passes[c];
		if (pass1) *pass1 = ctx->tune_pass1[c];
		return;
	}
	hj_radix_default(nr, bits, passes);
}
uint32_t hj_radix_bits_onepass(hj_ctx_t *ctx, uint64_t nr, uint32_t *variant) {
	const uint32_t c = hj_log2_ceil(nr ? nr : 1);
	if (variant) *variant = 0;
	// 8 bits are the ceiling: a one-pass scatter's runs get short above that
	// and its write-combining lines take more LDS than two workgroups per CU
	// leave (C4's 2048-join batch: 78 ms untuned at <= 8 bits against 108 ms
	// with the default's up to 9-10 bits; tuned 65-68 ms at <= 7-8 bits).
	// A tuned class may leave a single join unpartitioned (one table); in a
	// batch, whose join table is sized for ~2K-tuple partitions, a build side
	// above 2048 tuples then takes the default's bits instead.
	uint32_t b = 0, p = 0;
	hj_radix_default(nr, &b, &p);
	if (ctx && ctx->tuned && c < 40 && ctx->tune_bits1[c] != 0xff) {
		if (variant) *variant = ctx->tune_var1[c];
		if (ctx->tune_bits1[c] || nr <= 2048) b = ctx->tune_bits1[c];
	}
	return std::min(b, 8u);
}

static int radix_engine(hj_ctx_t *ctx, const hj_relation_t *R, const hj_relation_t *S, const hj_params_t *prm,
                        int tk, hj_result_t *out, uint32_t in_wv, const hj_relation_t *Rfb, const hj_relation_t *Sfb);

// The tuned configuration for |R| = nr when the caller leaves radix_bits and
// passes at 0: the skew-tuned one (hj_autotune_zipf) under
// HJ_F_SKEW_HANDLING when its class was tuned, else hj_radix_config's.
// *pscale: the probe-range scale the skew tuner chose (0 otherwise).
static void radix_choose(hj_ctx_t *ctx, uint64_t nr, const hj_params_t *prm, uint32_t *B, uint32_t *P, uint32_t *b1,
                         uint32_t *pscale) {
	const uint32_t c = hj_log2_ceil(nr ? nr : 1);
	*pscale = 0;
	if ((prm->flags & HJ_F_SKEW_HANDLING) && ctx && ctx->tuned_z && c < 40 && ctx->tune_bits_z[c] != 0xff) {
		*B = ctx->tune_bits_z[c];
		*P = ctx->tune_passes_z[c];
		*b1 = ctx->tune_pass1_z[c];
		*pscale = ctx->tune_pscale_z[c];
		return;
	}
	hj_radix_config(ctx, nr, B, P, b1);
}

// passes the engine takes for |R| = nr (the first lines of radix_engine)
static uint32_t radix_plan_passes(hj_ctx_t *ctx, uint64_t nr, const hj_params_t *prm) {
	uint32_t B = prm->radix_bits, P = prm->passes, tb1 = 0, ps = 0;
	if (B == 0 && P == 0) radix_choose(ctx, nr, prm, &B, &P, &tb1, &ps);
	if (P == 0) P = default_passes(B);
	if (B == 0) P = 0;
	if (P == 2 && B < 2) P = 1;
	return P;
}

// A radix join's result (count, checksum) depends on every tuple's 16-byte
// head {key, row_id} only, so for tuples wider than 16 bytes the partitions
// and the join move heads: at 64 bytes 64 + 8 + 16 + 8 B per tuple (narrow
// heads) instead of 5 x 64 (the reference carries whole tuples through both
// passes, store_a_tuple, memaccess.h:309-355).  Two passes read the heads in
// place in pass A; one pass or none copies them out first.
// Two radix passes address a tuple inside its relation with 32 bits (31 with
// the narrow-tuple mark).  Larger relations are joined in pieces of 2^30
// tuples: every pair (r, s) lies in exactly one (R piece, S piece) join and
// the count and checksum are sums, so the pieces' results add up to the
// join's.  An lp join keeps the first build row per key, which pieces of R
// would not see across pieces: such a join runs as nphj_lp, whose result is
// the same.
static const uint64_t kPiece = 1ull << 30;

int hj_run_radix(hj_ctx_t *ctx, const hj_relation_t *R, const hj_relation_t *S, const hj_params_t *prm, int tk,
                 hj_result_t *out) {
	if (tk == HJ_TK_HRO && R->tuple_size > 64)
		return hj_fail(ctx, HJ_EINVAL, "hro joins tuples of at most 64 bytes (one 512-bit block holds a tuple)");
	const uint32_t W = R->tuple_size, wv = W / 16;
	if ((R->tuple_num >= (1ull << 31) || S->tuple_num >= (1ull << 31)) &&
	    radix_plan_passes(ctx, R->tuple_num, prm) == 2) {
		if (ctx->emit_pairs)
			return hj_fail(ctx, HJ_EINVAL, "late-materialising radix joins take relations below 2^31 tuples");
		if (tk == HJ_TK_LP && R->tuple_num >= (1ull << 31)) {
			hj_params_t p2{};
			return hj_run_nphj_lp(ctx, R, S, &p2, out);
		}
		// the radix configuration of the whole build side, for every piece
		hj_params_t pp = *prm;
		if (pp.radix_bits == 0 && pp.passes == 0) {
			uint32_t tb1 = 0, ps = 0;
			radix_choose(ctx, R->tuple_num, &pp, &pp.radix_bits, &pp.passes, &tb1, &ps);
			if (!pp.radix_bits_pass1) pp.radix_bits_pass1 = tb1;
			if (!((pp.part_variant >> 20) & 3)) pp.part_variant |= ps << 20;
		}
		hj_result_t acc{};
		for (uint64_t r0 = 0; r0 < std::max<uint64_t>(R->tuple_num, 1); r0 += kPiece) {
			hj_relation_t rp = *R;
			rp.tuples = (void *)((const char *)R->tuples + r0 * W);
			rp.tuple_num = std::min(kPiece, R->tuple_num - r0);
			for (uint64_t s0 = 0; s0 < std::max<uint64_t>(S->tuple_num, 1); s0 += kPiece) {
				hj_relation_t sp = *S;
				sp.tuples = (void *)((const char *)S->tuples + s0 * W);
				sp.tuple_num = std::min(kPiece, S->tuple_num - s0);
				hj_result_t part{};
				const int rc = hj_run_radix(ctx, &rp, &sp, &pp, tk, &part);
				if (rc) return rc;
				acc.matched_cnt += part.matched_cnt;
				acc.check_sum += part.check_sum;
				acc.buildpart_time += part.buildpart_time;
				acc.probejoin_time += part.probejoin_time;
				acc.total_time += part.total_time;
				acc.algo_bytes += part.algo_bytes;
				acc.fallback |= part.fallback;
				acc.radix_bits = part.radix_bits;
				acc.passes = part.passes;
				acc.narrow_tuples += part.narrow_tuples;
				acc.positional_tuples += part.positional_tuples;
				acc.narrow = part.narrow;
			}
		}
		acc.n_ktimes = out->n_ktimes;
		*out = acc;
		return HJ_OK;
	}
	if (wv == 1) return radix_engine(ctx, R, S, prm, tk, out, 0, R, S);
	const uint64_t nr = R->tuple_num, ns = S->tuple_num;
	int rc;
	// (an lp join that emits row-id pairs may fall back to nphj_lp, whose
	// pair-emitting probe reads dense 16-byte heads: it takes the copy)
	if (radix_plan_passes(ctx, nr, prm) == 2 && !(ctx->emit_pairs && tk == HJ_TK_LP)) {
		const hj_relation_t hr{R->tuples, nr, 16, HJ_LOC_DEVICE}, hs{S->tuples, ns, 16, HJ_LOC_DEVICE};
		if ((rc = radix_engine(ctx, &hr, &hs, prm, tk, out, wv, R, S))) return rc;
		// pass A reads a burst of the wide tuple, not 16 bytes
		out->algo_bytes += (uint64_t)(std::min<uint32_t>(W, 64) - 16) * (nr + ns);
		return HJ_OK;
	}
	if ((rc = hj_reserve(ctx, ctx->keyrid[0], std::max<uint64_t>(nr, 1) * 16)) ||
	    (rc = hj_reserve(ctx, ctx->keyrid[1], std::max<uint64_t>(ns, 1) * 16)))
		return rc;
	// own events: the engine records ev[0..3] and the kernel pairs, a
	// late-materialising join around this one the EV_PTR ones
	const int e0 = hj_ctx::EV_HEAD;
	HJ_HIP(ctx, hipEventRecord(ctx->ev[e0], ctx->stream));
	hj_kt_begin(ctx, HJ_K_KEYRID);
	hj_launch_keyrid(ctx, R->tuples, nr, wv, ctx->keyrid[0].p);
	hj_launch_keyrid(ctx, S->tuples, ns, wv, ctx->keyrid[1].p);
	hj_kt_end(ctx);
	HJ_HIP(ctx, hipGetLastError());
	HJ_HIP(ctx, hipEventRecord(ctx->ev[e0 + 1], ctx->stream));
	const hj_relation_t kr{ctx->keyrid[0].p, nr, 16, HJ_LOC_DEVICE}, ks{ctx->keyrid[1].p, ns, 16, HJ_LOC_DEVICE};
	if ((rc = radix_engine(ctx, &kr, &ks, prm, tk, out, 0, &kr, &ks))) return rc;
	const double t_heads = hj_event_seconds(ctx, e0, e0 + 1);
	out->buildpart_time += t_heads;
	out->total_time += t_heads;
	out->algo_bytes += (uint64_t)(std::min<uint32_t>(W, 64) + 16) * (nr + ns);
	return HJ_OK;
}

// in_wv > 1: R and S hold tuples of in_wv 16-byte words and the engine
// works on their heads (pass A reads them in place; two passes only).  Rfb /
// Sfb: the caller's relations, for the exact recomputation of an lp join.
static int radix_engine(hj_ctx_t *ctx, const hj_relation_t *R, const hj_relation_t *S, const hj_params_t *prm,
                        int tk, hj_result_t *out, uint32_t in_wv, const hj_relation_t *Rfb, const hj_relation_t *Sfb) {
	const bool lp = tk == HJ_TK_LP;
	const uint32_t W = R->tuple_size, wv = W / 16;
	const uint64_t nr = R->tuple_num, ns = S->tuple_num, N = nr + ns;
	uint32_t B = prm->radix_bits, P = prm->passes, tuned_b1 = 0, tuned_ps = 0;
	if (B == 0 && P == 0) radix_choose(ctx, nr, prm, &B, &P, &tuned_b1, &tuned_ps);
	// probe tuples per work item: part_variant bits 20-21, else the skew tuner's
	const uint32_t pscale = ((prm->part_variant >> 20) & 3) ? ((prm->part_variant >> 20) & 3) : tuned_ps;
	auto scale_probe = [&](uint64_t v, uint64_t lo) -> uint64_t {
		switch (pscale) {
		case 1: return std::max<uint64_t>(lo, v / 2);
		case 2: return std::max<uint64_t>(lo, v / 4);
		case 3: return v * 2;
		default: return v;
		}
	};
	if (P == 0) P = default_passes(B);
	if (B == 0) P = 0;
	if (P == 2 && B < 2) P = 1;
	// 512-byte tuples are partitioned by the tile sorts only (a one-pass
	// scatter tile of them would not leave LDS for its write-combining
	// lines); more radix bits never change the result
	if (wv == 32 && P == 1) {
		B = std::max(B, 2u);
		P = 2;
	}
	// two passes: by default the first takes one bit more than the larger
	// half (its 4096-tuple tiles keep runs of 16 tuples for pass B to gather
	// and the second pass's 2048-tuple tiles then leave longer runs for the
	// join); the tuner or the caller (NUM_RADIX_BITS_PASS1) may choose
	// another split
	uint32_t b1 = P == 2 ? std::min(std::min((B + 1) / 2 + 1, B - 1), 12u) : (P ? B : 0);
	if (P == 2 && tuned_b1 && tuned_b1 < B) b1 = tuned_b1;
	if (P == 2 && (prm->part_variant >> 8 & 15) && (prm->part_variant >> 8 & 15) < B)
		b1 = prm->part_variant >> 8 & 15;  // first-pass width forced for tuning sweeps
	if (P == 2 && prm->radix_bits_pass1) {
		if (prm->radix_bits_pass1 >= B)
			return hj_fail(ctx, HJ_EINVAL, "radix_bits_pass1 %u must be below radix_bits %u", prm->radix_bits_pass1, B);
		b1 = prm->radix_bits_pass1;
	}
	const uint32_t b2 = B - b1;
	const uint32_t F1 = 1u << b1, F2 = 1u << b2, Q = 1u << B;
	if (b1 > 12 || (P == 2 && b2 > 12)) return hj_fail(ctx, HJ_EINVAL, "at most 12 radix bits per pass");
	if (P == 2 && (nr >> 31 || ns >> 31))  // hj_run_radix splits larger relations
		return hj_fail(ctx, HJ_EINVAL, "two radix passes take relations of fewer than 2^31 tuples");
	out->radix_bits = B;
	out->passes = P;
	hj_stats *st = (hj_stats *)ctx->stats.p;
	const uint32_t variant = prm->part_variant;
	// narrow tuples mark pass-A sources in bit 31 of a segment offset
	ctx->narrow_ok = wv == 1 && !(prm->flags & HJ_F_NO_NARROW) && nr < (1ull << 31) && ns < (1ull << 31);
	// positional words: row ids equal to relation positions + 1 (as every
	// relation of the reference's generator, src/gen.c:205, :244, :262) are
	// not stored by pass A; offsets below 2^30 leave two code bits in pass
	// B's staged sources
	ctx->dense_ok = ctx->narrow_ok && P == 2 && !(prm->flags & HJ_F_NO_POSITIONAL) && nr < (1ull << 30) &&
	                ns < (1ull << 30);
	const part_cfg pc = cfg_for(wv, variant);

	// chunk size: a multiple of the scatter tile, ~4 chunks per CU
	auto chunk_for = [&](uint64_t n) {
		uint64_t target = (uint64_t)ctx->num_cu * 4;
		uint64_t ch = (n + target - 1) / target;
		ch = (ch + pc.tile - 1) / pc.tile * pc.tile;
		if (ch < (uint64_t)pc.tile) ch = pc.tile;
		return ch;
	};
	// pass-B tile: one 32 KB sort tile, or for 16-byte tuples 4096 by 1024
	// threads (part_variant bits 18-19 = 1, tuning sweeps)
	const part_cfg pc2 = (wv == 1 && ((variant >> 18) & 3) == 1) ? part_cfg{1024, 4, 4096} : cfg_for(wv, 0);
	const part_cfg pcA = cfg_passA(wv, variant, b1, in_wv);
	// one pass: histogram + scatter over large chunks; two passes: pass A
	// sorts sort-tile-sized chunks in place
	const uint64_t ch1 = P == 2 ? (uint64_t)pcA.tile : chunk_for(N), ch2 = pc2.tile;
	const uint64_t chunks1 = (nr + ch1 - 1) / ch1 + (ns + ch1 - 1) / ch1;
	const uint32_t nseg2 = 2 * F1;
	const uint64_t chunks2_max = N / ch2 + nseg2 + 1;
	const uint64_t hist1_len = chunks1 * F1, hist2_len = 0;

	// scratch
	int rc;
	// slack: a gathering kernel re-reads the first tuple for an empty tile
	if (P >= 1 && (rc = hj_reserve(ctx, ctx->part[0], N * W + 4096))) return rc;
	if (P == 2 && (rc = hj_reserve(ctx, ctx->part[1], N * W + 4096))) return rc;
	const uint64_t hist_len = std::max(hist1_len, hist2_len) + 1;
	if ((rc = hj_reserve(ctx, ctx->hist, hist_len * 8))) return rc;
	if ((rc = hj_reserve(ctx, ctx->scan_tmp, ((hist_len + Q + nseg2) / SCAN_CH + 2) * 8 + 4096))) return rc;
	// meta: segs1[2] | segs2[nseg2] | cmap[max chunks] | pb1 | pairs | tiles | chunk starts
	const uint64_t cmap_n = std::max<uint64_t>(chunks1, chunks2_max);
	size_t off = 0;
	auto take = [&](size_t bytes) {
		size_t o = off;
		off += (bytes + 255) & ~(size_t)255;
		return o;
	};
	const size_t o_seg1 = take(2 * sizeof(hj_seg)), o_seg2 = take(nseg2 * sizeof(hj_seg));
	const size_t o_cmap = take(cmap_n * 4), o_pb1 = take(((uint64_t)2 * F1 + 1) * 8);
	const size_t o_pairs = take((size_t)Q * sizeof(hj_pair)), o_tiles = take(((size_t)Q + 1) * 8);
	const size_t o_cst = take(P == 2 ? chunks2_max * 8 : 8);
	const size_t o_runA = take(P == 2 ? chunks1 * (F1 + 1) * 2 : 8), o_cstA = take(P == 2 ? chunks1 * 8 : 8);
	const size_t o_begt = take(P == 2 ? (hist1_len + 1) * 2 : 8);
	const size_t o_gt = take(P == 2 ? chunks2_max * sizeof(hj_gtile) : 8);
	const size_t o_runt = take(P == 2 ? chunks2_max * (F2 + 1) * 2 : 8);
	const size_t o_tnar = take(P == 2 ? chunks1 : 8);
	// two passes: groups of 64 pass-A tiles per segment, their per-digit sums
	// (scanned), and the join's per-first-digit items
	const uint32_t GR = (uint32_t)(((nr + ch1 - 1) / ch1 + 63) / 64), GS = (uint32_t)(((ns + ch1 - 1) / ch1 + 63) / 64);
	const size_t o_gpart = take(P == 2 ? ((uint64_t)F1 * (GR + GS) + 1) * 4 : 8);
	const size_t o_dpre = take(P == 2 ? ((uint64_t)F1 + 1) * 8 : 8), o_dtiles = take(P == 2 ? (uint64_t)F1 * 4 : 8);
	if ((rc = hj_reserve(ctx, ctx->meta, off))) return rc;
	char *meta = (char *)ctx->meta.p;
	hj_seg *segs1 = (hj_seg *)(meta + o_seg1), *segs2 = (hj_seg *)(meta + o_seg2);
	uint32_t *cmap = (uint32_t *)(meta + o_cmap);
	uint64_t *pb1 = (uint64_t *)(meta + o_pb1);
	uint64_t *tiles = (uint64_t *)(meta + o_tiles);
	hj_pair *pairs = (hj_pair *)(meta + o_pairs);
	uint16_t *runsA = (uint16_t *)(meta + o_runA);
	uint64_t *cst = (uint64_t *)(meta + o_cst), *cstA = (uint64_t *)(meta + o_cstA);
	hj_gtile *gt = (hj_gtile *)(meta + o_gt);
	uint16_t *beg_t = (uint16_t *)(meta + o_begt), *run_t = (uint16_t *)(meta + o_runt);
	ctx->tile_narrow = (uint8_t *)(meta + o_tnar);
	uint32_t *gpart = (uint32_t *)(meta + o_gpart), *dtiles = (uint32_t *)(meta + o_dtiles);
	uint64_t *dpre = (uint64_t *)(meta + o_dpre);
	// probe tiles of the two-pass join: cpt pass-B tiles of S per item, one
	// item per second digit of a uniform partition
	uint32_t cpt = 0;
	if (P == 2) {
		const uint64_t avg_chunks = (ns / F1 + pc2.tile - 1) / pc2.tile;
		cpt = (uint32_t)scale_probe(std::max<uint64_t>(16, avg_chunks + avg_chunks / 4 + 2), 4);
	}
	uint64_t *hist = (uint64_t *)ctx->hist.p;
	// two passes: the count matrix holds 32-bit positions (a piece's R and S
	// have at most 2^31 tuples each, so every position fits)
	uint32_t *hist32 = (uint32_t *)ctx->hist.p;

	HJ_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
	hj_stats_reset(ctx);
	const void *Rb = R->tuples, *Sb = S->tuples;
	const uint64_t *pbF = nullptr;
	const size_t sldsA = (size_t)pcA.tile * wv * 16 + (2 * (size_t)F1 + 1) * 4 + 64 * 4;
	const size_t sldsB = gathersort_lds(pc2, wv, F2);
	if (P >= 1) {
		hipLaunchKernelGGL(k_seg_pair, dim3(std::max<unsigned>(1, (unsigned)((chunks1 + 255) / 256))), dim3(256), 0,
		                   ctx->stream, (const uint4 *)R->tuples, nr, (const uint4 *)S->tuples, ns, ch1, F1, segs1,
		                   cmap, st, in_wv);
		if (P == 1) {
			rc = launch_pass(ctx, wv, variant, (unsigned)chunks1, segs1, cmap, ch1, 0, b1, hist, hist1_len,
			                 ctx->part[0].p, st, (uint64_t *)ctx->scan_tmp.p);
			if (rc) return rc;
		} else {
			// pass A: every input tile sorted by the first digit in place; its
			// run tables give the histogram
			if (sldsA + 1024 > ctx->prop.sharedMemPerBlock || sldsB + 1024 > ctx->prop.sharedMemPerBlock)
				return hj_fail(ctx, HJ_EINVAL, "radix bits per pass too large for LDS");
			const unsigned gA = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(chunks1, (uint64_t)ctx->num_cu * 4));
			// kernel times: the two-pass regions are chained (one event
			// between two regions), the last one closing at ev[1]
			hj_kt_begin(ctx, HJ_K_TILESORT);
			launch_sortchunk(ctx, wv, pcA, gA, sldsA, segs1, cmap, 0, b1, ctx->part[0].p, runsA, cstA);
			// the count matrix: group sums, their scan, positions
			hj_kt_next(ctx, HJ_K_SCAN);
			const dim3 gg(std::max(1u, GR + GS), (F1 + 63) / 64);
			hipLaunchKernelGGL(k_runs_part, gg, dim3(256), 0, ctx->stream, runsA, segs1, F1, GR, GS, gpart);
			rc = scan_u32(ctx, gpart, (uint64_t)F1 * (GR + GS), (uint64_t *)ctx->scan_tmp.p);
			hipLaunchKernelGGL(k_runs_pos, gg, dim3(256), 0, ctx->stream, runsA, segs1, F1, GR, GS, gpart, hist32,
			                   beg_t);
			if (rc) return rc;
		}
		if (P == 1)
			hipLaunchKernelGGL(k_bounds<uint64_t>, dim3((unsigned)std::min<uint64_t>(1024, (2ull * F1 + 256) / 256)),
			                   dim3(256), 0, ctx->stream, segs1, 2u, F1, (const uint64_t *)hist, pb1);
		pbF = pb1;
		Rb = Sb = ctx->part[0].p;
		if (P == 2) {
			// second-pass segments, pass-B tile records and the join's work list
			hj_kt_next(ctx, HJ_K_PLAN);
			allow_lds(ctx, (const void *)k_segs2);
			hipLaunchKernelGGL(k_segs2, dim3(1), dim3(1024), (size_t)16 * F1, ctx->stream, segs1, gpart, F1, GR, GS,
			                   ch2, (const uint4 *)ctx->part[0].p, wv, b2, cpt, segs2, pb1, dpre, dtiles, st, hist32,
			                   hist1_len);
			// pass-B order's digit blocks: part_variant bits 22-25 = v > 0
			// take 2^(v-1) digits (tuning sweeps), else 8
			const uint32_t pb_dv = (variant >> 22) & 15;
			const uint32_t pb_dblk = pb_dv ? 1u << (pb_dv - 1) : 8u;
			allow_lds(ctx, (const void *)k_first_run);
			hipLaunchKernelGGL(k_first_run, dim3((unsigned)std::min<uint64_t>(4096, (chunks2_max + 255) / 256)),
			                   dim3(256), (size_t)24 * F1, ctx->stream, segs1, segs2, st, b1, (uint32_t)ch2, hist32,
			                   gt, b2, dpre, dtiles, pairs, tiles, std::min(pb_dblk, F1));
			// pass B: each first-level partition's tiles gathered from the
			// pass-A runs and sorted by the second digit
			// two workgroups per CU: with three the tiles in flight on an XCD
			// touch about as many lines as its 4 MB L2 holds, and lines two
			// runs share are evicted before the second read (pass B 1.06 vs
			// 1.28 ms at C2); part_variant bits 14-15 (tuning sweeps) take
			// 1: all resident, 2: one per CU
			const uint32_t gsel = (variant >> 14) & 3;
			const unsigned sgrid = gsel == 1 ? ~0u : (unsigned)ctx->num_cu * (gsel == 2 ? 1u : 2u);
			hj_kt_next(ctx, HJ_K_GATHERSORT);
			launch_sortchunk_g(ctx, wv, pc2, sgrid, sldsB, gt, hist32, beg_t, b1, b2, ctx->part[0].p, ctx->part[1].p, run_t,
			                   (uint32_t)chunks2_max, cst, (uint32_t)pcA.tile);
			Rb = Sb = ctx->part[1].p;
		}
	}
	HJ_HIP(ctx, hipGetLastError());
	HJ_HIP(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
	if (P == 2) hj_kt_close_at(ctx, 1);  // the pass-B region

	// join: tile the probe side so a skewed partition spreads over workgroups
	const uint64_t avg_s = (ns >> B) + 1;
	uint64_t s_tile = scale_probe(std::max<uint64_t>(8192, 2 * avg_s), 2048);
	const uint64_t avg_r = (nr + Q - 1) / Q;
	const size_t lds_max = std::min<size_t>(ctx->prop.sharedMemPerBlock, 160 * 1024) - join_static_lds(tk, P == 2) - 256;
	uint32_t cap = (uint32_t)std::min<uint64_t>(65535, std::max<uint64_t>(256, avg_r + avg_r / 4 + 64));
	while (cap > 256u && hj_tab_lds(tk, cap, wv) > lds_max) cap -= 256;
	const size_t jlds = hj_tab_lds(tk, cap, wv);
	if (jlds > lds_max) return hj_fail(ctx, HJ_EINVAL, "join table of %u tuples does not fit LDS", cap);
	const size_t jall = jlds + join_static_lds(tk, P == 2);
	const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(4, (size_t)(160 * 1024) / jall));
	const unsigned jgrid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)ctx->num_cu * per_cu, (uint64_t)Q * 64));
	// few partitions against a large probe side (a small build relation):
	// tiles of twice the average partition would leave most of the grid
	// idle, so cut the probe side into about two items per workgroup
	if (P <= 1 && (uint64_t)Q * ((avg_s + s_tile - 1) / s_tile) < 2ull * jgrid)
		s_tile = std::max<uint64_t>(8192, ns / (2ull * jgrid) + 1);
	if (P <= 1) {
		hj_kt_begin(ctx, HJ_K_PLAN);
		hipLaunchKernelGGL(k_plan, dim3((Q + 255) / 256), dim3(256), 0, ctx->stream, pbF, P, Q, b1, F2, nr, ns, s_tile,
		                   pairs, tiles);
		rc = scan_u64(ctx, tiles, Q, (uint64_t *)ctx->scan_tmp.p);
		hj_kt_end(ctx);
		if (rc) return rc;
	}
	if (P == 2) {
		hj_kt_open_at(ctx, HJ_K_JOIN, 1);
		launch_join2(ctx, wv, tk, jgrid, jlds, Rb, cst, run_t, (uint32_t)chunks2_max, b1, pairs, tiles, Q, B, cpt, cap);
	} else {
		hj_kt_begin(ctx, HJ_K_JOIN);
		launch_join(ctx, wv, tk, jgrid, jlds, Rb, Sb, pairs, tiles, Q, B, s_tile, cap);
		hj_kt_end(ctx);
	}
	HJ_HIP(ctx, hipGetLastError());
	HJ_HIP(ctx, hipEventRecord(ctx->ev[2], ctx->stream));
	if (P == 2) hj_kt_close_at(ctx, 2);
	hj_stats h;
	if ((rc = hj_stats_read(ctx, &h))) return rc;
	out->buildpart_time = hj_event_seconds(ctx, 0, 1);
	out->probejoin_time = hj_event_seconds(ctx, 1, 2);
	out->total_time = hj_event_seconds(ctx, 0, 2);
	// each pass reads and writes whole tuples; the join reads heads only, one
	// 64-byte HBM burst of a wider tuple
	out->algo_bytes = (uint64_t)(2 * P) * W * N + (uint64_t)std::min(W, 64u) * N;
	if (P == 2 && ctx->narrow_ok) {
		// pass A reads 16 B and writes 8 or 16 B per tuple; pass B reads
		// that and writes 8 B when no tile was wide; the join reads pass B's
		const uint64_t nt = h.narrow_tuples, nd = h.dense_tuples, a_out = 4 * nd + 8 * (nt - nd) + 16 * (N - nt);
		out->narrow_tuples = nt;
		out->positional_tuples = nd;
		out->narrow = !(h.flags & HJ_DF_WIDE);
		const uint64_t b_out = (out->narrow ? 8 : 16) * N;
		out->algo_bytes = 16 * N + a_out + a_out + b_out + b_out;
	}
	if (lp && (h.flags & (HJ_DF_DUPKEY | HJ_DF_OVERSIZE))) {
		// Exact path: linear probing keeps only the first build tuple per key
		// (src/algo/phj.c:655-666 with a stable partition = smallest R row),
		// which is what nphj_lp computes with its duplicate fix-up (here with
		// a multiplicative hash, whatever the keys' low bits).
		hj_result_t r2{};
		rc = hj_run_lp_exact(ctx, Rfb, Sfb, &r2);
		if (rc) return rc;
		out->matched_cnt = r2.matched_cnt;
		out->check_sum = r2.check_sum;
		// the re-run is part of this join's work: its time joins the probe phase
		out->probejoin_time += r2.total_time;
		out->total_time += r2.total_time;
		out->fallback |= (h.flags & HJ_DF_DUPKEY) ? HJ_FB_LP_DUPLICATES : 0;
		out->fallback |= (h.flags & HJ_DF_OVERSIZE) ? HJ_FB_LP_OVERSIZE : 0;
		return HJ_OK;
	}
	out->matched_cnt = h.matched;
	out->check_sum = h.checksum;
	return HJ_OK;
}

// ======================================================================
// Batched joins: many independent joins in one captured HIP graph.
//
// Planning (host, once per batch shape): each join gets a size class -- a
// build side that fits one LDS table is not partitioned; larger ones get the
// radix bits of their class (one pass).  All partitioned joins' R and S
// become segments of ONE multi-split launch (per-segment digit width), then
// a planner turns every (join, partition) into probe tiles and one
// persistent kernel joins all tiles of all joins, adding each tile's count
// and checksum into that join's result slot.  Integer sums mod 2^64 are
// order-independent, so each join's result equals its stand-alone result.
// ======================================================================

struct hj_bjoin {
	const uint4 *R;
	const uint4 *S;
	uint64_t nr, ns;
	uint32_t bits;       // 0: not partitioned
	uint32_t seg;        // index of its first segment (R, then S; S only if R is resident)
	uint32_t pair_base;  // first pair index of this join
	uint32_t pad;
	uint64_t s_tile;     // probe tuples per work item
	const uint64_t *rpb; // resident, already partitioned build side: bounds ...
	const uint4 *rpart;  // ... and tuples (null when R is partitioned here)
};

struct hj_bpair {
	const uint4 *rb, *sb;
	uint64_t r0, r1, s0, s1;
	uint32_t join, shift;
	uint32_t rn, sn;  // build / probe side held as 8-byte tuples (narrow pass output)
};

__global__ void k_batch_reset(hj_stats *st, uint64_t nchunks, unsigned long long *res, uint32_t njoins,
                              unsigned int *jflags) {
	const uint32_t stride = gridDim.x * blockDim.x;
	for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < njoins; i += stride) {
		res[2 * i] = 0;
		res[2 * i + 1] = 0;
		jflags[i] = 0;
	}
	if (blockIdx.x == 0 && threadIdx.x < 8) st->xq[threadIdx.x] = 0;  // gathered batches' queues
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		st->work_next = 0;
		st->flags = 0;
		st->total_chunks = nchunks;
		st->narrow_tuples = 0;
		st->dense_tuples = 0;
	}
}

// pb[seg.pb_base + p] = first position of digit p, pb[seg.pb_base + F] = end
__global__ void k_bounds_b(const hj_seg *__restrict__ segs, uint32_t nseg, const uint64_t *__restrict__ scan,
                           uint64_t *__restrict__ pb) {
	for (uint32_t s = blockIdx.x; s < nseg; s += gridDim.x) {
		const hj_seg g = segs[s];
		const uint32_t F = 1u << g.bits;
		for (uint32_t p = threadIdx.x; p <= F; p += blockDim.x)
			pb[g.pb_base + p] = p == F ? g.out_start + g.len
			                           : (g.nchunks ? scan[g.hist_base + (uint64_t)p * g.nchunks] : g.out_start);
	}
}

// narrow_ok: the pass that wrote `part` could write 8-byte tuples; whether
// it did is in st->flags (no HJ_DF_WIDE).  Inputs and resident partitions
// are always 16-byte.
__global__ void k_plan_b(const hj_bjoin *__restrict__ joins, const uint32_t *__restrict__ pair_join, uint32_t npairs,
                         const hj_seg *__restrict__ segs, const uint64_t *__restrict__ pb, const uint4 *part,
                         hj_bpair *__restrict__ pairs, uint64_t *__restrict__ tiles, const hj_stats *__restrict__ st,
                         bool narrow_ok) {
	const uint32_t pn = narrow_ok && !(st->flags & HJ_DF_WIDE) ? 1u : 0u;
	for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < npairs; i += gridDim.x * blockDim.x) {
		const uint32_t j = pair_join[i];
		const hj_bjoin J = joins[j];
		hj_bpair pr;
		pr.join = j;
		pr.shift = J.bits;
		pr.rn = pr.sn = 0;
		if (J.bits == 0) {
			pr.rb = J.R;
			pr.sb = J.S;
			pr.r0 = 0;
			pr.r1 = J.nr;
			pr.s0 = 0;
			pr.s1 = J.ns;
		} else if (J.rpb) {
			const uint32_t q = i - J.pair_base;
			const uint64_t sb = segs[J.seg].pb_base + q;
			pr.rb = J.rpart;
			pr.sb = part;
			pr.sn = pn;
			pr.r0 = *hj_g(&J.rpb[q]);
			pr.r1 = *hj_g(&J.rpb[q + 1]);
			pr.s0 = pb[sb];
			pr.s1 = pb[sb + 1];
		} else {
			const uint32_t q = i - J.pair_base;
			const uint64_t rb = segs[J.seg].pb_base + q, sb = segs[J.seg + 1].pb_base + q;
			pr.rb = pr.sb = part;
			pr.rn = pr.sn = pn;
			pr.r0 = pb[rb];
			pr.r1 = pb[rb + 1];
			pr.s0 = pb[sb];
			pr.s1 = pb[sb + 1];
		}
		pairs[i] = pr;
		const uint64_t rl = pr.r1 - pr.r0, sl = pr.s1 - pr.s0;
		tiles[i] = (rl && sl) ? (sl + J.s_tile - 1) / J.s_tile : 0;
	}
}

// Persistent bc join over (join, partition, tile) items; per-item results
// are wave-reduced and added to res[2*join], res[2*join+1].  A partition
// larger than the table is joined in several fills.
template <int WV, int NT>
__global__ __launch_bounds__(NT) void k_join_b(const hj_bpair *__restrict__ pairs,
                                               const uint64_t *__restrict__ tprefix, uint32_t npairs,
                                               const hj_bjoin *__restrict__ joins, uint32_t cap, hj_stats *st,
                                               unsigned long long *__restrict__ res) {
	extern __shared__ uint64_t lds_u64[];
	__shared__ uint64_t s_item[4];
	uint64_t *keys = lds_u64;
	uint64_t *pays = keys + cap;
	uint32_t *heads = (uint32_t *)(pays + cap);
	const uint32_t nbmax = 1u << (32 - __clz(cap - 1));
	uint16_t *next = (uint16_t *)(heads + nbmax);
	const uint64_t total = tprefix[npairs];
	const bool lane0 = (threadIdx.x & (HJ_WAVE - 1)) == 0;

	for (;;) {
		if (threadIdx.x == 0) {
			const uint64_t w = atomicAdd(&st->work_next, 1ULL);
			uint32_t lo = 0, hi = npairs;
			while (hi - lo > 1) {
				const uint32_t mid = (lo + hi) >> 1;
				if (tprefix[mid] <= w) lo = mid; else hi = mid;
			}
			s_item[0] = w;
			s_item[1] = lo;
			s_item[2] = w < total ? w - tprefix[lo] : 0;
		}
		__syncthreads();
		const uint64_t w = s_item[0];
		const uint32_t q = (uint32_t)s_item[1];
		const uint64_t tix = s_item[2];
		__syncthreads();
		if (w >= total) break;
		const hj_bpair pr = pairs[q];
		const uint32_t B = pr.shift;
		const uint64_t s_tile = joins[pr.join].s_tile;
		const uint64_t sb = pr.s0 + tix * s_tile, se = min(pr.s1, sb + s_tile);
		unsigned long long cnt = 0, sum = 0;
		for (uint64_t rc = pr.r0; rc < pr.r1; rc += cap) {
			const uint32_t m = (uint32_t)min((uint64_t)cap, pr.r1 - rc);
			const uint32_t nb = m <= 1 ? 1u : 1u << (32 - __clz(m - 1));
			const uint32_t hmask = nb - 1;
			for (uint32_t i = threadIdx.x; i < nb; i += NT) heads[i] = 0;
			__syncthreads();
			// four build tuples per thread and step, loads issued together
			// (unconditional, clamped to the last tuple); T = uint2 for a
			// narrow side (8-byte index), uint4 otherwise (WV words a tuple)
			auto build = [&](auto tag) {
				using T = decltype(tag);
				constexpr uint32_t SW = sizeof(T) == 8 ? 1 : WV;
				const T *rb = (const T *)pr.rb;
				for (uint32_t i0 = threadIdx.x; i0 < m; i0 += 4 * NT) {
					T t[4];
#pragma unroll
					for (int u = 0; u < 4; u++) {
						const T *a = &rb[(rc + min(i0 + u * NT, m - 1)) * SW];
						if constexpr (sizeof(T) == 8) t[u] = hj_load8(a);
						else t[u] = hj_load(a);
					}
#pragma unroll
					for (int u = 0; u < 4; u++) {
						const uint32_t i = i0 + u * NT;
						if (i >= m) continue;
						const uint64_t key = hj_key_of(t[u]);
						keys[i] = key;
						pays[i] = hj_pay_of(t[u]);
						next[i] = (uint16_t)atomicExch(&heads[(uint32_t)(key >> B) & hmask], i + 1);
					}
				}
			};
			if (WV == 1 && pr.rn) build(uint2{});
			else build(uint4{});
			__syncthreads();
			// probe tuples U per thread and step; the next step's loads are
			// issued (unconditional, clamped to the last tuple) before the
			// current step is probed
			constexpr int U = 4;
			auto run_probe = [&](auto tag) {
			using T = decltype(tag);
			constexpr uint32_t SW = sizeof(T) == 8 ? 1 : WV;
			const T *sbase = (const T *)pr.sb;
			auto load = [&](T (&t)[U], uint64_t j0) {
#pragma unroll
				for (int u = 0; u < U; u++) {
					const T *a = &sbase[min(j0 + (uint64_t)u * NT, se - 1) * SW];
					if constexpr (sizeof(T) == 8) t[u] = hj_load_stream8(a);
					else t[u] = hj_load_stream(a);
				}
			};
			auto probe = [&](const T (&t)[U], uint64_t j0) {
#pragma unroll
				for (int u = 0; u < U; u++) {
					if (j0 + (uint64_t)u * NT >= se) continue;
					const uint64_t key = hj_key_of(t[u]), sp = hj_pay_of(t[u]);
					for (uint32_t hit = heads[(uint32_t)(key >> B) & hmask]; hit; hit = next[hit - 1])
						if (keys[hit - 1] == key) {
							cnt++;
							sum += pays[hit - 1] + sp;
						}
				}
			};
			// the same number of steps in every lane (sb < se), loads clamped
			// and never under a lane condition, so each probe waits only for
			// its own buffer while the other one's loads are in flight
			const uint64_t step = (uint64_t)U * NT;
			const uint64_t nsteps = (se - sb + step - 1) / step;
			T ta[U], tb[U];
			load(ta, sb + threadIdx.x);
			for (uint64_t k = 0; k < nsteps; k += 2) {
				const uint64_t j0 = sb + k * step + threadIdx.x;
				load(tb, j0 + step);
				probe(ta, j0);
				load(ta, j0 + 2 * step);
				probe(tb, j0 + step);  // past the end: every lane skips
			}
			};
			if (WV == 1 && pr.sn) run_probe(uint2{});
			else run_probe(uint4{});
			__syncthreads();
		}
		cnt = hj_wave_sum(cnt);
		sum = hj_wave_sum(sum);
		if (lane0 && cnt) {
			atomicAdd(&res[2 * pr.join], cnt);
			atomicAdd(&res[2 * pr.join + 1], sum);
		}
	}
}

// ---------------------------------------------------------------------------
// Gathered batches (16-byte tuples, bc and lp): no histogram pass and no scatter.
// Every relation of every join is cut into 8192-tuple tiles that pass A
// (k_sortchunk<..., true>) counting-sorts by the join's own digit in place,
// storing a tile as 4-byte positional words, 8-byte {key, row_id} or 16-byte
// tuples, with a run table per tile.  k_run_transpose_b makes the run tables
// digit-major per segment, and one persistent k_join_g joins partition
// (join, digit d) by gathering run d of every tile of the join's R (the LDS
// table) and of a range of its S tiles (the probes).  A partition's lines
// that neighbouring digits share are read by workgroups of one XCD at about
// the same time (items in (join, digit) order from one queue per XCD).

static constexpr uint32_t kGTile = 8192;  // pass-A tile of gathered batches (1024 threads x 8)
static constexpr uint32_t kGTileBits = 13;

// One 64 x 64 block of a segment's run table ([tile][digit], F + 1 entries
// a tile, from segs[seg].hist_base) written digit-major into run_t (row d of
// the segment at hist_base + d * nchunks).
struct hj_tblk {
	uint32_t seg, c0, p0, pad;
};
__global__ __launch_bounds__(256) void k_run_transpose_b(const uint16_t *__restrict__ run, const hj_seg *__restrict__ segs,
                                                         const hj_tblk *__restrict__ blk, uint16_t *__restrict__ run_t) {
	__shared__ uint16_t t[64][66];
	const hj_tblk b = blk[blockIdx.x];
	const hj_seg g = segs[b.seg];
	const uint32_t F = 1u << g.bits, nc = g.nchunks;
	for (uint32_t e = threadIdx.x; e < 64 * 64; e += 256) {
		const uint32_t ci = e / 64, pj = e % 64, c = b.c0 + ci, p = b.p0 + pj;
		t[ci][pj] = (c < nc && p <= F) ? run[g.hist_base + (uint64_t)c * (F + 1) + p] : (uint16_t)0;
	}
	__syncthreads();
	for (uint32_t e = threadIdx.x; e < 64 * 64; e += 256) {
		const uint32_t pj = e / 64, ci = e % 64, c = b.c0 + ci, p = b.p0 + pj;
		if (c < nc && p <= F) run_t[g.hist_base + (uint64_t)p * nc + c] = t[ci][pj];
	}
}

// A join of a gathered launch: per side its digit-major run rows (row d of
// a side with nc tiles at rt + d * nc) and tile descriptors.  A resident
// build relation keeps its own (serving), so these are pointers.
struct hj_gjoin {
	const uint16_t *rtR, *rtS;
	const uint64_t *dR, *dS;
	uint32_t ncR, ncS;    // tiles of R, of S
	uint32_t bits, cpt;   // radix bits; S tiles per work item
};
// A work item: join, digit (12 bits) | probe-tile range index << 12.
static constexpr uint32_t kGItemDigitBits = 12, kGItemDigits = 1u << kGItemDigitBits;
struct hj_gitem {
	uint32_t join, dt;
};

// a run element of any layout (hj_load16a4 / hj_load8a4: dword alignment)
__device__ __forceinline__ void hj_gload(uint4 &v, const char *p) { v = hj_load16a4c(p); }
__device__ __forceinline__ void hj_gload(uint2 &v, const char *p) { v = hj_load8a4c(p); }

// A gathered tuple: v its 16 loaded bytes, pz its tile's segment position
// (a multiple of kGTile) | layout code.  kh = key >> bits (all keys of a
// partition share the low `bits` bits, so kh decides equality and hashes as
// bc_join's (key >> B)); rid = row id.
__device__ __forceinline__ void hj_gdecode(const uint4 &v, uint32_t pz, uint32_t bits, uint64_t &kh, uint64_t &rid) {
	const uint32_t code = pz & 3u;
	const uint64_t k16 = (uint64_t)v.x | (uint64_t)v.y << 32, r16 = (uint64_t)v.z | (uint64_t)v.w << 32;
	const uint64_t kp = v.x >> kGTileBits, rp = (uint64_t)((pz & ~(kGTile - 1)) + (v.x & (kGTile - 1)) + 1);
	const uint64_t k8 = (uint64_t)(v.x >> bits), r8 = v.y;
	kh = code == 2 ? kp : code == 1 ? k8 : (k16 >> bits);
	rid = code == 2 ? rp : code == 1 ? r8 : r16;
}
// the same from the first 8 bytes, when no tile is 16-byte (layout 1 or 2):
// kh and rid fit 32 bits
__device__ __forceinline__ void hj_gdecode(const uint2 &v, uint32_t pz, uint32_t bits, uint32_t &kh, uint32_t &rid) {
	const bool pos = (pz & 3u) == 2;
	kh = pos ? v.x >> kGTileBits : v.x >> bits;
	rid = pos ? (pz & ~(kGTile - 1)) + (v.x & (kGTile - 1)) + 1 : v.y;
}

// Persistent join of a gathered batch.  Work item = (pair (join, digit),
// probe-tile index): S tiles [tix * cpt, (tix + 1) * cpt) of that join.
// NARROW (no 16-byte tile anywhere: every kh and rid fits 32 bits) keeps
// the LDS table as u64 {kh, rid}; otherwise as two u64 arrays.  Both are
// launched; the one that does not match the pass-A flag returns at once.
// LDS: table (kv | [rid] | heads u32[np2(cap)] | next u16[cap]).
template <int NT, bool NARROW>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void k_join_g(
    const hj_gjoin *__restrict__ joins, const hj_gitem *__restrict__ items, const uint64_t *__restrict__ qb,
    uint32_t cap, hj_stats *st, unsigned long long *__restrict__ res, unsigned int *__restrict__ jflags, bool lp) {
	if (NARROW == ((st->flags & HJ_DF_WIDE) != 0)) return;
	// loaded bytes per tuple: 8 when no tile is 16-byte (an 8-byte load of a
	// 4-byte word reads 4 bytes of its tile beyond it), else 16
	using VT = typename std::conditional<NARROW, uint2, uint4>::type;
	using KT = typename std::conditional<NARROW, uint32_t, uint64_t>::type;
	constexpr uint32_t NW = NT / HJ_WAVE;
	constexpr uint32_t RU = 4;               // tuples per lane per step
	constexpr uint32_t STEP = HJ_WAVE * RU;  // tuples per wave per step
	extern __shared__ uint64_t lds_u64[];
	__shared__ uint32_t sred[NW + 1];
	__shared__ uint64_t s_item[4];
	__shared__ uint64_t ent_a[NT + 1];  // staged run k: address of its first tuple | layout code
	__shared__ uint2 ent_o[NT + 1];     // {flat offset, segment position of its tile}; ent_o[NT].x = total
	// the probe side is staged by each wave alone: 64 tiles' runs at a time
	__shared__ uint64_t went_a[NW][HJ_WAVE];
	__shared__ uint2 went_o[NW][HJ_WAVE + 1];
	const uint32_t nb = hj_np2u(cap), hmask = nb - 1;
	uint64_t *kv = lds_u64;
	uint64_t *rv = kv + cap;  // wide table: row ids
	uint32_t *heads = (uint32_t *)(kv + (NARROW ? cap : 2 * cap));
	uint16_t *next = (uint16_t *)(heads + nb);
	const uint32_t lane = threadIdx.x & (HJ_WAVE - 1), wid = threadIdx.x / HJ_WAVE;
	const uint64_t total = qb[8];
	uint32_t cur = blockIdx.x & 7;  // this workgroup's queue (its XCD)
	// thread 0 claims the next item while the current one runs: kpend is the
	// position taken in queue `cur` (its value is waited for only when the
	// next item starts)
	uint64_t kpend = 0;
	if (threadIdx.x == 0) kpend = atomicAdd(&st->xq[cur], 1ULL);

	// runs of digit d of tiles [t0, t1) (t1 - t0 <= NT) of a side (run rows
	// at rt, nc tiles, descriptors dsc); the first run starts `skip` tuples in
	auto stage = [&](const uint16_t *rt, const uint64_t *dsc, uint32_t nc, uint32_t t0, uint32_t t1, uint32_t d,
	                 uint32_t skip) -> uint32_t {
		const uint32_t t = t0 + threadIdx.x;
		uint32_t len = 0;
		uint64_t addr = 0;
		if (t < t1) {
			uint32_t a = hj_g(rt)[(uint64_t)d * nc + t];
			len = hj_g(rt)[(uint64_t)(d + 1) * nc + t] - a;
			if (threadIdx.x == 0) {
				a += skip;
				len -= skip;
			}
			const uint64_t desc = hj_g(dsc)[t];
			const uint32_t code = (uint32_t)desc & 3u;
			addr = ((desc & ~3ull) + (uint64_t)a * (16u >> code)) | code;
		}
		uint32_t tot;
		const uint32_t ex = hj_block_exscan<uint32_t>(len, sred, &tot);
		ent_a[threadIdx.x] = addr;
		ent_o[threadIdx.x] = make_uint2(ex, t * kGTile);
		if (threadIdx.x == NT - 1) ent_o[NT] = make_uint2(tot, 0u);
		return tot;
	};
	struct cursor_t {
		uint32_t k, off, noff, pz;
		uint64_t a;
	};
	auto seek = [&](uint32_t f) -> cursor_t {
		uint32_t lo = 0, hi = NT;
		while (hi - lo > 1) {
			const uint32_t mid = (lo + hi) >> 1;
			if (ent_o[mid].x <= f) lo = mid; else hi = mid;
		}
		const uint64_t a = ent_a[lo];
		return cursor_t{lo, ent_o[lo].x, ent_o[lo + 1].x, ent_o[lo].y | ((uint32_t)a & 3u), a};
	};
	// cursor at flat index 0 of entry arrays EA / EO
	auto first = [&](const uint64_t *EA, const uint2 *EO) -> cursor_t {
		const uint64_t a = EA[0];
		return cursor_t{0, EO[0].x, EO[1].x, EO[0].y | ((uint32_t)a & 3u), a};
	};
	// flat indices f0 + u*64 + lane, clamped to fe - 1 (every lane issues RU
	// loads, none under a lane condition); the cursor only moves forward
	auto load_step = [&](uint32_t f0, uint32_t fe, cursor_t &cs, VT (&v)[RU], uint32_t (&pz)[RU], const uint64_t *EA,
	                     const uint2 *EO) {
#pragma unroll
		for (uint32_t u = 0; u < RU; u++) {
			const uint32_t f = min(f0 + u * HJ_WAVE + lane, fe - 1);
			while (f >= cs.noff) {
				cs.k++;
				cs.off = cs.noff;
				cs.noff = EO[cs.k + 1].x;
				cs.a = EA[cs.k];
				cs.pz = EO[cs.k].y | ((uint32_t)cs.a & 3u);
			}
			const uint32_t code = (uint32_t)cs.a & 3u;
			hj_gload(v[u], (const char *)(uintptr_t)(cs.a & ~3ull) + (uint64_t)(f - cs.off) * (16u >> code));
			pz[u] = cs.pz;
		}
	};
	unsigned long long cnt = 0, sum = 0;
	auto probe_step = [&](uint32_t f0, uint32_t fe, const VT (&v)[RU], const uint32_t (&pz)[RU], uint32_t bits) {
		uint32_t c[RU];
		KT kh[RU], rid[RU];
#pragma unroll
		for (uint32_t u = 0; u < RU; u++) {
			hj_gdecode(v[u], pz[u], bits, kh[u], rid[u]);
			c[u] = f0 + u * HJ_WAVE + lane < fe ? heads[(uint32_t)kh[u] & hmask] : 0u;
		}
		bool any = true;
		while (any) {
			any = false;
#pragma unroll
			for (uint32_t u = 0; u < RU; u++) {
				if (!c[u]) continue;
				if (NARROW) {
					const uint64_t e = kv[c[u] - 1];
					if ((uint32_t)e == (uint32_t)kh[u]) {
						cnt++;
						sum += (e >> 32) + (uint64_t)rid[u];
					}
				} else if (kv[c[u] - 1] == kh[u]) {
					cnt++;
					sum += rv[c[u] - 1] + rid[u];
				}
				c[u] = next[c[u] - 1];
				any |= c[u] != 0;
			}
		}
	};

	for (;;) {
		if (threadIdx.x == 0) {
			uint64_t w = total;
			for (uint32_t tried = 0; tried < 8;) {
				const uint64_t w0 = qb[cur], w1 = qb[cur + 1];
				if (w0 + kpend < w1) {
					w = w0 + kpend;
					break;
				}
				if (++tried == 8) break;
				cur = (cur + 1) & 7;
				kpend = atomicAdd(&st->xq[cur], 1ULL);
			}
			s_item[0] = w;
			if (w < total) kpend = atomicAdd(&st->xq[cur], 1ULL);  // the next item's claim
		}
		__syncthreads();
		const uint64_t w = s_item[0];
		__syncthreads();
		if (w >= total) break;
		const hj_gitem it = items[w];
		const hj_gjoin J = joins[it.join];
		const uint32_t d = it.dt & (kGItemDigits - 1), tix = it.dt >> kGItemDigitBits;
		const uint32_t sb = tix * J.cpt, se = min(J.ncS, sb + J.cpt);
		uint32_t rc = 0;    // next R tile
		uint32_t roff = 0;  // tuples of tile rc's run already taken by an earlier fill
		while (rc < J.ncR) {
			// ---- fill the table with up to `cap` build tuples
			for (uint32_t i = threadIdx.x; i < nb; i += NT) heads[i] = 0;
			__syncthreads();
			uint32_t m = 0;
			for (;;) {
				const uint32_t c1 = min(J.ncR, rc + NT);
				const uint32_t n = c1 - rc;
				const uint32_t tot = stage(J.rtR, J.dR, J.ncR, rc, c1, d, roff);
				__syncthreads();
				const uint32_t T = min(tot, cap - m);
				// runs taken whole form a prefix; the next one may be taken in part
				const uint32_t nfull = __syncthreads_count(threadIdx.x < n && ent_o[threadIdx.x + 1].x <= T);
				const uint32_t roff_new = nfull < n ? (nfull == 0 ? roff : 0) + (T - ent_o[nfull].x) : 0;
				const uint32_t fb = (uint32_t)((uint64_t)T * wid / NW), fe = (uint32_t)((uint64_t)T * (wid + 1) / NW);
				if (fb < fe) {
					cursor_t cs = seek(fb);
					for (uint32_t f0 = fb; f0 < fe; f0 += STEP) {
						VT v[RU];
						uint32_t pz[RU];
						load_step(f0, fe, cs, v, pz, ent_a, ent_o);
#pragma unroll
						for (uint32_t u = 0; u < RU; u++) {
							const uint32_t f = f0 + u * HJ_WAVE + lane;
							if (f >= fe) continue;
							KT kh, rid;
							hj_gdecode(v[u], pz[u], J.bits, kh, rid);
							const uint32_t i = m + f;
							if (NARROW) {
								kv[i] = (uint64_t)(uint32_t)kh | (uint64_t)rid << 32;
							} else {
								kv[i] = kh;
								rv[i] = rid;
							}
							next[i] = (uint16_t)atomicExch(&heads[(uint32_t)kh & hmask], i + 1);
						}
					}
				}
				m += T;
				rc += nfull;
				roff = roff_new;
				__syncthreads();
				if (m >= cap || rc >= J.ncR) break;
			}
			// lp (first build tuple per key): without duplicate build keys its
			// result is bc's; a duplicate in the table, or a partition needing a
			// second fill (a duplicate could span fills), flags the join for the
			// host's exact path, which replaces its result
			if (lp) {
				bool dup = false;
				for (uint32_t i = threadIdx.x; i < m; i += NT) {
					const uint64_t ki = NARROW ? (uint64_t)(uint32_t)kv[i] : kv[i];
					for (uint32_t c = heads[(uint32_t)ki & hmask]; c && !dup; c = next[c - 1])
						dup = c - 1 != i && (NARROW ? (uint64_t)(uint32_t)kv[c - 1] : kv[c - 1]) == ki;
				}
				const bool more = rc < J.ncR;
				if (__syncthreads_or(dup) && threadIdx.x == 0) atomicOr(&jflags[it.join], HJ_DF_DUPKEY);
				if (more && threadIdx.x == 0) atomicOr(&jflags[it.join], HJ_DF_OVERSIZE);
			}
			// ---- probe S tiles [sb, se): each wave stages and probes chunks of
			// cw tiles (wave w: chunks w, w + NW, ...) with no block barrier,
			// the next chunk's run bounds and tile descriptors loaded while the
			// current one is probed
			if (m) {
				const uint32_t nt = se - sb;
				const uint32_t cw = min((uint32_t)HJ_WAVE, max(1u, (nt + NW - 1) / NW));
				uint64_t *EA = went_a[wid];
				uint2 *EO = went_o[wid];
				uint32_t ma = 0, me = 0;
				uint64_t mdesc = 0;
				// chunk [t0, t0 + cw) ∩ [.., se): lane l's tile, clamped (every lane loads)
				auto meta_load = [&](uint32_t t0) {
					const uint32_t t = min(t0 + lane, se - 1);  // t0 may be past the end
					ma = hj_g(J.rtS)[(uint64_t)d * J.ncS + t];
					me = hj_g(J.rtS)[(uint64_t)(d + 1) * J.ncS + t];
					mdesc = hj_g(J.dS)[t];
				};
				auto meta_commit = [&](uint32_t t0) -> uint32_t {
					const uint32_t t = t0 + lane;
					const uint32_t len = (lane < cw && t < se) ? me - ma : 0u;
					uint32_t incl = len;
#pragma unroll
					for (int o = 1; o < HJ_WAVE; o <<= 1) {
						const uint32_t x = __shfl_up(incl, o, HJ_WAVE);
						if ((int)lane >= o) incl += x;
					}
					const uint32_t tot = __shfl(incl, HJ_WAVE - 1, HJ_WAVE);
					const uint32_t code = (uint32_t)mdesc & 3u;
					EA[lane] = ((mdesc & ~3ull) + (uint64_t)ma * (16u >> code)) | code;
					EO[lane] = make_uint2(incl - len, t * kGTile);
					if (lane == HJ_WAVE - 1) EO[HJ_WAVE] = make_uint2(tot, 0u);
					__builtin_amdgcn_wave_barrier();
					return tot;
				};
				// loads are never under a condition (indices are clamped): a load
				// in a conditional block makes the compiler wait at the merge for
				// every load issued in it, which would serialise the prefetch
				uint32_t t0 = sb + wid * cw;
				meta_load(t0);
				while (t0 < se) {
					const uint32_t T = meta_commit(t0);
					const uint32_t tn = t0 + NW * cw;
					meta_load(tn);
					if (T) {
						cursor_t cs = first(EA, EO);
						VT va[RU], vb[RU];
						uint32_t pa[RU], pb[RU];
						load_step(0, T, cs, va, pa, EA, EO);
						for (uint32_t f0 = 0; f0 < T; f0 += 2 * STEP) {
							load_step(f0 + STEP, T, cs, vb, pb, EA, EO);
							probe_step(f0, T, va, pa, J.bits);
							load_step(f0 + 2 * STEP, T, cs, va, pa, EA, EO);
							probe_step(f0 + STEP, T, vb, pb, J.bits);  // lanes past T skip
						}
					}
					__builtin_amdgcn_wave_barrier();
					t0 = tn;
				}
			}
			__syncthreads();  // every wave is done with this fill's table
		}
		// this item's matches go to its join's slot (one atomic pair per wave)
		cnt = hj_wave_sum(cnt);
		sum = hj_wave_sum(sum);
		if (lane == 0 && cnt) {
			atomicAdd(&res[2 * it.join], cnt);
			atomicAdd(&res[2 * it.join + 1], sum);
		}
		cnt = sum = 0;
		__syncthreads();
	}
}

static constexpr int kGJoinNT = 1024;  // one join workgroup per CU (its table fills the LDS)

// LDS bytes of k_join_g's table for `cap` build tuples
static size_t join_g_lds(uint32_t cap, bool narrow) {
	return (size_t)cap * (narrow ? 10 : 18) + (size_t)hj_np2u(cap) * 4 + 8;
}

// Threads of a batch-join workgroup: its LDS table (about 64 KB for 2048
// build tuples) lets two share a CU, so 512 threads keep 16 waves per CU.
static constexpr int kJoinBNT = 512;

template <int WV>
static void launch_join_b_t(hj_ctx_t *ctx, unsigned grid, size_t lds, const hj_bpair *pairs, const uint64_t *tiles,
                            uint32_t npairs, const hj_bjoin *joins, uint32_t cap, hj_stats *st, unsigned long long *res) {
	hipLaunchKernelGGL((k_join_b<WV, kJoinBNT>), dim3(grid), dim3(kJoinBNT), lds, ctx->stream, pairs, tiles, npairs, joins,
	                   cap, st, res);
}

static void launch_join_b(hj_ctx_t *ctx, uint32_t wv, unsigned grid, size_t lds, const hj_bpair *pairs,
                          const uint64_t *tiles, uint32_t npairs, const hj_bjoin *joins, uint32_t cap, hj_stats *st,
                          unsigned long long *res) {
#define JB(W) launch_join_b_t<W>(ctx, grid, lds, pairs, tiles, npairs, joins, cap, st, res)
	switch (wv) {
	case 1: JB(1); break;
	case 2: JB(2); break;
	case 4: JB(4); break;
	case 8: JB(8); break;
	default: JB(16); break;
	}
#undef JB
}

struct hj_batch {
	hj_ctx_t *ctx = nullptr;
	bool lp = false;
	uint32_t njoins = 0, wv = 1, W = 16;
	std::vector<hj_relation_t> R, S;
	hj_buf part, hist, scan_tmp, meta, res, stats;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	hj_batch_info_t info{};
	// device pointers into meta
	unsigned long long *d_res = nullptr;
	unsigned int *d_jflags = nullptr;
	// algorithmic bytes: gathered batches count from the tile layouts pass A
	// chose; scattered ones from these tuple counts and the narrow flag
	bool gathered = false;
	uint64_t part_tuples = 0, unpart_tuples = 0;
	// the batch's launch sequence on ctx->stream (captured into the graph;
	// a batch created with HJ_F_KERNEL_TIMES runs it directly, each kernel
	// between two events)
	std::function<void()> launch;
	bool timed = false;
	// lp batch run join by join through the exact first-match path (no graph)
	bool exact = false;
};

static void batch_free(hj_batch *b) {
	if (b->exec) (void)hipGraphExecDestroy(b->exec);
	if (b->graph) (void)hipGraphDestroy(b->graph);
	hj_buf *bufs[] = {&b->part, &b->hist, &b->scan_tmp, &b->meta, &b->res, &b->stats};
	for (hj_buf *x : bufs)
		if (x->p) (void)hipFree(x->p);
	if (b->e0) (void)hipEventDestroy(b->e0);
	if (b->e1) (void)hipEventDestroy(b->e1);
	delete b;
}

static const void *join_b_kernel(uint32_t wv) {
	switch (wv) {
	case 1: return (const void *)k_join_b<1, kJoinBNT>;
	case 2: return (const void *)k_join_b<2, kJoinBNT>;
	case 4: return (const void *)k_join_b<4, kJoinBNT>;
	case 8: return (const void *)k_join_b<8, kJoinBNT>;
	default: return (const void *)k_join_b<16, kJoinBNT>;
	}
}

// Radix bits of a join in a gathered batch: at most 6144 build tuples per
// partition on average (the join's LDS table holds 8192), so that a run of
// one digit in an 8192-tuple tile is 8192 * 3072..6144 / |R| tuples long;
// build sides up to 6144 tuples are one partition.
static constexpr uint64_t kGPart = 6144;
static uint32_t gbatch_bits(uint64_t nr) {
	if (nr <= kGPart) return 0;
	return std::min<uint32_t>(10, hj_log2_ceil((nr + kGPart - 1) / kGPart));
}
// ... or the autotuner's choice for the size class, timed through one-join
// gathered batches
static uint32_t gbatch_bits(const hj_ctx_t *ctx, uint64_t nr) {
	const uint32_t c = hj_log2_ceil(nr ? nr : 1);
	if (ctx && ctx->tuned && c < 40 && ctx->tune_gbits[c] != 0xff) return std::min<uint32_t>(10, ctx->tune_gbits[c]);
	return gbatch_bits(nr);
}

// Whether a batch runs gathered: 16-byte joins (lp ones flag duplicate build
// keys and multi-fill partitions for the exact path) whose build sides
// hold at most 2^22 tuples (10 radix bits keep their partitions within one
// LDS table) and whose probe sides hold fewer than 2^30 (positions and row
// ids of 4-byte words fit 32 bits), unless the caller asks for the scatter
// path.
static bool gbatch_eligible(bool lp, uint32_t W, uint32_t njoins, const hj_relation_t *R, const hj_relation_t *S,
                            const hj_params_t *params) {
	(void)lp;  // lp joins flag duplicate build keys for the exact path, as scattered batches do
	if (W != 16 || (params && (params->flags & HJ_F_BATCH_SCATTER))) return false;
	for (uint32_t j = 0; j < njoins; j++)
		if (R[j].tuple_num > (1ull << 22) || S[j].tuple_num >= (1ull << 30)) return false;
	return true;
}

// A build relation kept in gathered layout (serving's resident builds):
// tiles sorted by the digit of `bits`, digit-major run rows, descriptors.
struct hj_gres {
	const uint16_t *rt = nullptr;
	const uint64_t *desc = nullptr;
	uint32_t nc = 0, bits = 0;
};

// The plan of a gathered launch: segments (pass-A input), run-table
// transposition blocks, joins and work items, and where each lives in one
// contiguous device region:
//   part | run | run_t | desc | (uploaded:) segs | cmap | joins | items | qb | blocks
struct hj_gplan {
	std::vector<hj_seg> segs;
	std::vector<uint32_t> cmap;
	std::vector<hj_tblk> blocks;
	std::vector<hj_gjoin> joins;
	std::vector<int32_t> jseg;  // R segment of a join (-1: resident), S segment is the next one unless resident
	std::vector<int32_t> jsseg; // S segment
	std::vector<hj_gitem> items;
	uint64_t qb[9] = {};
	uint64_t part_n = 0, runs_n = 0, tuples = 0;
	uint32_t maxbits = 0, n_part = 0, n_unpart = 0;
	// byte offsets
	size_t o_part = 0, o_run = 0, o_runt = 0, o_desc = 0, o_up = 0, o_segs = 0, o_cmap = 0, o_joins = 0, o_items = 0,
	       o_qb = 0, o_blk = 0, bytes = 0;
};

static uint32_t gplan_seg(hj_gplan &P, const hj_relation_t &rel, uint32_t bits) {
	hj_seg g{};
	g.in = (const uint4 *)rel.tuples;
	g.len = rel.tuple_num;
	g.out_start = P.part_n;
	g.bits = bits;
	g.chunk_base = (uint32_t)P.cmap.size();
	g.nchunks = (uint32_t)((g.len + kGTile - 1) / kGTile);
	g.hist_base = P.runs_n;
	const uint32_t F = 1u << bits;
	P.runs_n += (uint64_t)g.nchunks * (F + 1);
	P.part_n += g.len;
	P.maxbits = std::max(P.maxbits, bits);
	const uint32_t si = (uint32_t)P.segs.size();
	for (uint32_t c = 0; c < g.nchunks; c++) P.cmap.push_back(si);
	for (uint32_t c0 = 0; c0 < g.nchunks; c0 += 64)
		for (uint32_t p0 = 0; p0 <= F; p0 += 64) P.blocks.push_back(hj_tblk{si, c0, p0, 0});
	P.segs.push_back(g);
	return si;
}

// Add join (R ⋈ S); `res` non-null: R is resident in gathered layout.
static void gplan_add(const hj_ctx_t *ctx, hj_gplan &P, const hj_relation_t &R, const hj_relation_t &S,
                      const hj_gres *res) {
	hj_gjoin J{};
	const uint32_t bits = res ? res->bits : gbatch_bits(ctx, R.tuple_num);
	int32_t rs = -1;
	if (res) {
		J.rtR = res->rt;
		J.dR = res->desc;
		J.ncR = res->nc;
	} else {
		rs = (int32_t)gplan_seg(P, R, bits);
		J.ncR = P.segs[rs].nchunks;
	}
	const int32_t ss = (int32_t)gplan_seg(P, S, bits);
	J.ncS = P.segs[ss].nchunks;
	J.bits = bits;
	J.cpt = std::max<uint32_t>(1, J.ncS);
	P.tuples += R.tuple_num + S.tuple_num;
	if (bits) P.n_part++;
	else P.n_unpart++;
	P.joins.push_back(J);
	P.jseg.push_back(rs);
	P.jsseg.push_back(ss);
}

// Work items and the region layout.  Items: one per (join, digit) with the
// whole probe side; when that leaves fewer than four items per workgroup of
// the join grid (one per CU), probe sides are cut into tile ranges of at
// least four times the build side's tiles.  Items in (join, digit, range)
// order; eight queues of consecutive items, one per XCD.
static void gplan_finish(hj_gplan &P, uint32_t grid) {
	uint64_t n_whole = 0;
	for (const hj_gjoin &J : P.joins)
		if (J.ncR && J.ncS) n_whole += 1ull << J.bits;
	if (n_whole && n_whole < 4ull * grid) {
		const uint64_t k = (4ull * grid + n_whole - 1) / n_whole;
		for (hj_gjoin &J : P.joins) {
			const uint32_t cpt = (uint32_t)((J.ncS + k - 1) / k);
			J.cpt = std::max<uint32_t>(std::max<uint32_t>(1, 4 * J.ncR), cpt);
		}
	}
	for (uint32_t j = 0; j < P.joins.size(); j++) {
		const hj_gjoin &J = P.joins[j];
		if (!J.ncR || !J.ncS) continue;
		const uint32_t per = (J.ncS + J.cpt - 1) / J.cpt;
		for (uint32_t d = 0; d < (1u << J.bits); d++)
			for (uint32_t t = 0; t < per; t++) P.items.push_back(hj_gitem{j, d | t << kGItemDigitBits});
	}
	for (int x = 0; x <= 8; x++) P.qb[x] = (uint64_t)P.items.size() * x / 8;
	size_t off = 0;
	auto take = [&](size_t bytes) {
		size_t o = off;
		off += (bytes + 255) & ~(size_t)255;
		return o;
	};
	P.o_part = take(P.part_n * 16 + 64);
	P.o_run = take(P.runs_n * 2 + 2);
	P.o_runt = take(P.runs_n * 2 + 2);
	P.o_desc = take(P.cmap.size() * 8 + 8);
	P.o_up = off;
	P.o_segs = take(P.segs.size() * sizeof(hj_seg) + 8);
	P.o_cmap = take(P.cmap.size() * 4 + 4);
	P.o_joins = take(P.joins.size() * sizeof(hj_gjoin) + 8);
	P.o_items = take(P.items.size() * sizeof(hj_gitem) + 8);
	P.o_qb = take(sizeof(P.qb));
	P.o_blk = take(P.blocks.size() * sizeof(hj_tblk) + 16);
	P.bytes = off;
}

// Device pointers of every planned side, given the region's base; then the
// host image of the uploaded part [o_up, bytes) into `img`.
static void gplan_place(hj_gplan &P, char *base, unsigned char *img) {
	const uint16_t *runt = (const uint16_t *)(base + P.o_runt);
	const uint64_t *desc = (const uint64_t *)(base + P.o_desc);
	for (uint32_t j = 0; j < P.joins.size(); j++) {
		hj_gjoin &J = P.joins[j];
		if (P.jseg[j] >= 0) {
			const hj_seg &g = P.segs[P.jseg[j]];
			J.rtR = runt + g.hist_base;
			J.dR = desc + g.chunk_base;
		}
		const hj_seg &g = P.segs[P.jsseg[j]];
		J.rtS = runt + g.hist_base;
		J.dS = desc + g.chunk_base;
	}
	auto put = [&](size_t o, const void *src, size_t n) {
		if (n) memcpy(img + (o - P.o_up), src, n);
	};
	put(P.o_segs, P.segs.data(), P.segs.size() * sizeof(hj_seg));
	put(P.o_cmap, P.cmap.data(), P.cmap.size() * 4);
	put(P.o_joins, P.joins.data(), P.joins.size() * sizeof(hj_gjoin));
	put(P.o_items, P.items.data(), P.items.size() * sizeof(hj_gitem));
	put(P.o_qb, P.qb, sizeof(P.qb));
	put(P.o_blk, P.blocks.data(), P.blocks.size() * sizeof(hj_tblk));
}

// Everything a gathered launch needs on the device (plain values, so a batch
// can keep it in its launcher).
struct hj_glaunch {
	char *base = nullptr;
	uint64_t nchunks = 0, nblk = 0, nitems = 0;
	uint32_t njoins = 0, maxbits = 0, capn = 0, capw = 0;
	unsigned gA = 1, gn = 1, gw = 1;
	bool lp = false;  // first-match joins: flag duplicate build keys and multi-fill partitions
	size_t sldsA = 0, ldsn = 0, ldsw = 0;
	size_t o_part, o_run, o_runt, o_desc, o_segs, o_cmap, o_joins, o_items, o_qb, o_blk;
};

// Kernel attributes and grids of gathered launches, queried once per context
// (a server launches many small batches; each query is a runtime call).
// Called with ctx->mu held.
static int gplan_setup(hj_ctx_t *ctx) {
	if (ctx->g_ready) return HJ_OK;
	const void *kn = (const void *)k_join_g<kGJoinNT, true>, *kw = (const void *)k_join_g<kGJoinNT, false>;
	auto kA = k_sortchunk<1, 1024, kGTile / 1024, true>;
	// LDS tables: one workgroup of 1024 threads per CU holds the largest
	// table that fits (8192 tuples narrow); larger partitions join in fills
	const size_t lds_wg = std::min<size_t>(ctx->prop.sharedMemPerBlock, 160 * 1024) - 1024;
	auto fit_cap = [&](bool narrow, const void *kfn) {
		uint32_t c = 65535;
		while (c > 256 && join_g_lds(c, narrow) + static_lds(kfn) > lds_wg) c = c > 8192 ? hj_np2u(c) / 2 : c - 256;
		return c;
	};
	ctx->g_capn = fit_cap(true, kn);
	ctx->g_capw = fit_cap(false, kw);
	ctx->g_ldsn = join_g_lds(ctx->g_capn, true);
	ctx->g_ldsw = join_g_lds(ctx->g_capw, false);
	if (ctx->g_ldsw + static_lds(kw) > ctx->prop.sharedMemPerBlock || ctx->g_ldsn + static_lds(kn) > ctx->prop.sharedMemPerBlock)
		return hj_fail(ctx, HJ_EINVAL, "batch LDS table too big");
	allow_lds(ctx, (const void *)kA);
	allow_lds(ctx, kn);
	allow_lds(ctx, kw);
	for (uint32_t b = 0; b <= 10; b++)
		ctx->g_gA[b] = resident_grid(ctx, (const void *)kA, 1024, (size_t)kGTile * 16 + (2 * ((size_t)1 << b) + 1) * 4 + 64 * 4);
	ctx->g_gn = resident_grid(ctx, kn, kGJoinNT, ctx->g_ldsn);
	ctx->g_gw = resident_grid(ctx, kw, kGJoinNT, ctx->g_ldsw);
	ctx->g_ready = true;
	return HJ_OK;
}

static int gplan_prepare(hj_ctx_t *ctx, const hj_gplan &P, char *base, hj_glaunch &L, bool lp = false) {
	int rc;
	if ((rc = gplan_setup(ctx))) return rc;
	L.base = base;
	L.lp = lp;
	L.capn = ctx->g_capn;
	L.capw = ctx->g_capw;
	L.ldsn = ctx->g_ldsn;
	L.ldsw = ctx->g_ldsw;
	if (P.maxbits > 10) return hj_fail(ctx, HJ_EINVAL, "batch radix bits too big");
	L.sldsA = (size_t)kGTile * 16 + (2 * ((size_t)1 << P.maxbits) + 1) * 4 + 64 * 4;
	if (L.sldsA + 1024 > ctx->prop.sharedMemPerBlock) return hj_fail(ctx, HJ_EINVAL, "batch radix bits too big");
	if (P.blocks.size() >= (1ull << 31)) return hj_fail(ctx, HJ_EINVAL, "batch too large");
	L.nchunks = P.cmap.size();
	L.nblk = P.blocks.size();
	L.nitems = P.items.size();
	L.njoins = (uint32_t)P.joins.size();
	L.maxbits = P.maxbits;
	L.gA = std::max(1u, std::min<unsigned>((unsigned)std::min<uint64_t>(std::max<uint64_t>(1, L.nchunks), (uint64_t)ctx->num_cu * 4),
	                                       ctx->g_gA[P.maxbits]));
	L.gn = ctx->g_gn;
	L.gw = ctx->g_gw;
	L.o_part = P.o_part;
	L.o_run = P.o_run;
	L.o_runt = P.o_runt;
	L.o_desc = P.o_desc;
	L.o_segs = P.o_segs;
	L.o_cmap = P.o_cmap;
	L.o_joins = P.o_joins;
	L.o_items = P.o_items;
	L.o_qb = P.o_qb;
	L.o_blk = P.o_blk;
	return HJ_OK;
}

// The launch sequence on `strm` (per-kernel events when ctx->kernel_times).
static void gplan_launch(hj_ctx_t *ctx, hipStream_t strm, const hj_glaunch &L, hj_stats *st, unsigned long long *res,
                         unsigned int *fl) {
	char *B = L.base;
	hj_kt_begin(ctx, HJ_K_FILL);
	hipLaunchKernelGGL(k_batch_reset, dim3(std::max<uint32_t>(1, std::min<uint32_t>(1024, (L.njoins + 255) / 256))),
	                   dim3(256), 0, strm, st, L.nchunks, res, L.njoins, fl);
	hj_kt_end(ctx);
	if (L.nchunks) {
		hj_kt_begin(ctx, HJ_K_TILESORT);
		hipLaunchKernelGGL((k_sortchunk<1, 1024, kGTile / 1024, true>), dim3(L.gA), dim3(1024), L.sldsA, strm,
		                   (const hj_seg *)(B + L.o_segs), (const uint32_t *)(B + L.o_cmap), st, 0u, L.maxbits,
		                   (uint4 *)(B + L.o_part), (uint16_t *)(B + L.o_run), (uint64_t *)(B + L.o_desc), true, true,
		                   (uint8_t *)nullptr);
		hj_kt_end(ctx);
		hj_kt_begin(ctx, HJ_K_PLAN);
		hipLaunchKernelGGL(k_run_transpose_b, dim3((unsigned)L.nblk), dim3(256), 0, strm, (const uint16_t *)(B + L.o_run),
		                   (const hj_seg *)(B + L.o_segs), (const hj_tblk *)(B + L.o_blk), (uint16_t *)(B + L.o_runt));
		hj_kt_end(ctx);
	}
	if (L.nitems) {
		// both table layouts; the one that does not match pass A's flag returns
		const hj_gjoin *J = (const hj_gjoin *)(B + L.o_joins);
		const hj_gitem *I = (const hj_gitem *)(B + L.o_items);
		const uint64_t *Q = (const uint64_t *)(B + L.o_qb);
		hj_kt_begin(ctx, HJ_K_JOIN);
		hipLaunchKernelGGL((k_join_g<kGJoinNT, true>), dim3(L.gn), dim3(kGJoinNT), L.ldsn, strm, J, I, Q, L.capn, st, res,
		                   fl, L.lp);
		hipLaunchKernelGGL((k_join_g<kGJoinNT, false>), dim3(L.gw), dim3(kGJoinNT), L.ldsw, strm, J, I, Q, L.capw, st,
		                   res, fl, L.lp);
		hj_kt_end(ctx);
	}
}

// Plan and capture a gathered batch into b (whose R, S, njoins are set).
// Every relation holds fewer than 2^30 tuples, so pass A may store 4- and
// 8-byte tiles everywhere.
static int gbatch_create(hj_ctx_t *ctx, hj_batch *b) {
	const uint32_t njoins = b->njoins;
	hj_gplan P;
	for (uint32_t j = 0; j < njoins; j++) gplan_add(ctx, P, b->R[j], b->S[j], nullptr);
	gplan_finish(P, (uint32_t)ctx->num_cu);
	b->info.njoins = njoins;
	b->info.tuples = P.tuples;
	b->info.work_items = P.items.size();
	b->info.n_partitioned = P.n_part;
	b->info.n_unpartitioned = P.n_unpart;
	b->info.gathered = 1;
	b->gathered = true;
	b->part_tuples = P.tuples;
	int rc;
	if ((rc = hj_reserve(ctx, b->part, P.bytes))) return rc;
	if ((rc = hj_reserve(ctx, b->stats, sizeof(hj_stats)))) return rc;
	if ((rc = hj_reserve(ctx, b->res, (size_t)njoins * 20 + 64))) return rc;
	std::vector<unsigned char> img(P.bytes - P.o_up);
	gplan_place(P, (char *)b->part.p, img.data());
	if (!img.empty() && hipMemcpy((char *)b->part.p + P.o_up, img.data(), img.size(), hipMemcpyHostToDevice) != hipSuccess)
		return hj_fail(ctx, HJ_EHIP, "batch upload");
	b->d_res = (unsigned long long *)b->res.p;
	b->d_jflags = (unsigned int *)((char *)b->res.p + (size_t)njoins * 16);
	b->info.scratch_bytes = b->part.cap + b->res.cap + b->stats.cap;
	hj_glaunch L;
	if ((rc = gplan_prepare(ctx, P, (char *)b->part.p, L, b->lp))) return rc;
	hj_stats *d_st = (hj_stats *)b->stats.p;
	unsigned long long *d_res = b->d_res;
	unsigned int *d_fl = b->d_jflags;
	hipStream_t strm = ctx->stream;
	b->launch = [=]() { gplan_launch(ctx, strm, L, d_st, d_res, d_fl); };
	if (hipEventCreate(&b->e0) != hipSuccess || hipEventCreate(&b->e1) != hipSuccess) {
		(void)hipGetLastError();
		return hj_fail(ctx, HJ_EHIP, "event create");
	}
	if (b->timed) return HJ_OK;  // runs without a graph
	const bool kt = ctx->kernel_times;
	ctx->kernel_times = false;
	hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
	if (e != hipSuccess) {
		ctx->kernel_times = kt;
		return hj_fail(ctx, HJ_EHIP, "begin capture: %s", hipGetErrorString(e));
	}
	b->launch();
	hipGraph_t g = nullptr;
	e = hipStreamEndCapture(ctx->stream, &g);
	ctx->kernel_times = kt;
	if (e != hipSuccess) {
		if (g) (void)hipGraphDestroy(g);
		return hj_fail(ctx, HJ_EHIP, "end capture: %s", hipGetErrorString(e));
	}
	b->graph = g;
	e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
	if (e != hipSuccess) return hj_fail(ctx, HJ_EHIP, "graph instantiate: %s", hipGetErrorString(e));
	return HJ_OK;
}

extern "C" int hj_batch_create(hj_ctx_t *ctx, const char *algo, uint32_t njoins, const hj_relation_t *R,
                               const hj_relation_t *S, const hj_params_t *params, hj_batch_t **out) {
	if (!ctx || !algo || !out || (njoins && (!R || !S))) return hj_fail(ctx, HJ_EINVAL, "null argument");
	*out = nullptr;
	bool radix = false;
	const bool lp = hj_algo_is_lp(algo, &radix);
	if (!radix) return hj_fail(ctx, HJ_EINVAL, "batches run the radix joins (phj_radix_bc/lp), not '%s'", algo);
	if (params && (params->radix_bits || params->passes))
		return hj_fail(ctx, HJ_EINVAL, "batches choose radix bits per join: pass radix_bits = passes = 0");
	std::lock_guard<std::mutex> lk(ctx->mu);
	HJ_HIP(ctx, hipSetDevice(ctx->device));
	const uint32_t W = njoins ? R[0].tuple_size : 16;
	for (uint32_t j = 0; j < njoins; j++) {
		if (R[j].tuple_size != W || S[j].tuple_size != W)
			return hj_fail(ctx, HJ_EINVAL, "join %u: every relation of a batch needs the same tuple size", j);
		if (R[j].location != HJ_LOC_DEVICE || S[j].location != HJ_LOC_DEVICE)
			return hj_fail(ctx, HJ_EINVAL, "join %u: batch relations must be device resident", j);
		if ((R[j].tuple_num && !R[j].tuples) || (S[j].tuple_num && !S[j].tuples))
			return hj_fail(ctx, HJ_EINVAL, "join %u: null tuples", j);
	}
	if (W != 16 && W != 32 && W != 64 && W != 128 && W != 256) return hj_fail(ctx, HJ_EINVAL, "tuple size %u", W);
	hj_batch *b = new hj_batch();
	b->ctx = ctx;
	b->lp = lp;
	b->njoins = njoins;
	b->W = W;
	b->wv = W / 16;
	b->R.assign(R, R + njoins);
	b->S.assign(S, S + njoins);
	if (gbatch_eligible(lp, W, njoins, R, S, params)) {
		b->timed = params && (params->flags & HJ_F_KERNEL_TIMES);
		const int grc = gbatch_create(ctx, b);
		if (grc != HJ_OK) {
			batch_free(b);
			return grc;
		}
		*out = b;
		return HJ_OK;
	}
	if (lp) {
		// lp batches that do not run gathered (wider tuples, build sides above
		// 2^22, or HJ_F_BATCH_SCATTER) run every join through the exact
		// first-match path when the batch runs
		b->exact = true;
		b->info.njoins = njoins;
		for (uint32_t j = 0; j < njoins; j++) b->info.tuples += R[j].tuple_num + S[j].tuple_num;
		b->info.n_unpartitioned = njoins;
		*out = b;
		return HJ_OK;
	}
	const uint32_t wv = b->wv;
	// tile shape and write-combining line (part_variant bits 0-6): the
	// caller's, else the tuned shape of the size class holding most of the
	// batch's partitioned tuples (one scatter launch serves every join)
	uint32_t variant = params ? (params->part_variant & 0x7f) : 0;
	if (!variant && ctx->tuned) {
		uint64_t per_class[40] = {};
		for (uint32_t j = 0; j < njoins; j++)
			if (R[j].tuple_num > 2048) per_class[hj_log2_ceil(R[j].tuple_num) % 40] += R[j].tuple_num + S[j].tuple_num;
		uint32_t top = 0;
		for (uint32_t c = 1; c < 40; c++)
			if (per_class[c] > per_class[top]) top = c;
		if (per_class[top] && ctx->tune_bits1[top] != 0xff) variant = ctx->tune_var1[top];
	}
	const part_cfg pc = cfg_for(wv, variant);

	// ---- plan on the host
	const uint32_t small_max = 2048;  // build sides up to this size are joined unpartitioned
	std::vector<hj_bjoin> joins(njoins);
	std::vector<hj_seg> segs;
	std::vector<uint32_t> pair_join;
	uint64_t part_n = 0, tuples = 0;
	uint32_t maxbits = 0, pb_total = 0;
	uint64_t max_rq = 1;
	for (uint32_t j = 0; j < njoins; j++) {
		hj_bjoin &J = joins[j];
		J.rpb = nullptr;
		J.rpart = nullptr;
		J.R = (const uint4 *)R[j].tuples;
		J.S = (const uint4 *)S[j].tuples;
		J.nr = R[j].tuple_num;
		J.ns = S[j].tuple_num;
		tuples += J.nr + J.ns;
		uint32_t bits = 0;
		if (J.nr > small_max) bits = hj_radix_bits_onepass(ctx, J.nr);  // one pass in a batch
		J.bits = bits;
		J.seg = ~0u;
		J.pair_base = (uint32_t)pair_join.size();
		const uint32_t Q = 1u << bits;
		for (uint32_t q = 0; q < Q; q++) pair_join.push_back(j);
		J.s_tile = std::max<uint64_t>(8192, 2 * ((J.ns >> bits) + 1));
		max_rq = std::max<uint64_t>(max_rq, (J.nr >> bits) + 1);
		if (bits) {
			J.seg = (uint32_t)segs.size();
			maxbits = std::max(maxbits, bits);
			for (int side = 0; side < 2; side++) {
				hj_seg g{};
				g.in = side ? J.S : J.R;
				g.len = side ? J.ns : J.nr;
				g.out_start = part_n;
				g.bits = bits;
				g.pb_base = pb_total;
				pb_total += Q + 1;
				part_n += g.len;
				segs.push_back(g);
			}
			b->part_tuples += J.nr + J.ns;
		} else {
			b->info.n_unpartitioned++;
			b->unpart_tuples += J.nr + J.ns;
		}
	}
	b->info.n_partitioned = njoins - b->info.n_unpartitioned;
	b->info.njoins = njoins;
	b->info.tuples = tuples;
	// chunks: a multiple of the scatter tile, ~4 per CU over the whole batch
	uint64_t ch = (part_n + (uint64_t)ctx->num_cu * 4 - 1) / ((uint64_t)ctx->num_cu * 4);
	ch = std::max<uint64_t>(pc.tile, (ch + pc.tile - 1) / pc.tile * pc.tile);
	std::vector<uint32_t> cmap;
	uint64_t hist_len = 0;
	for (uint32_t s = 0; s < segs.size(); s++) {
		hj_seg &g = segs[s];
		g.chunk_base = (uint32_t)cmap.size();
		g.nchunks = (uint32_t)((g.len + ch - 1) / ch);
		g.hist_base = hist_len;
		hist_len += (uint64_t)g.nchunks << g.bits;
		for (uint32_t c = 0; c < g.nchunks; c++) cmap.push_back(s);
	}
	const uint64_t nchunks = cmap.size();
	const uint32_t npairs = (uint32_t)pair_join.size();
	const uint32_t nseg = (uint32_t)segs.size();
	const size_t slds = scatter_lds(pc, wv, 1u << maxbits, pass_line(ctx, pc, wv, 1u << maxbits, variant));
	// LDS table: expected partition size with headroom, at most 3400 tuples
	// so that two workgroups share a CU (the batch's largest partition would
	// otherwise size every workgroup's table: C4 77 -> 68 ms); a larger
	// partition is joined in several fills
	const size_t lds_max = std::min<size_t>(ctx->prop.sharedMemPerBlock, 160 * 1024) - 1024;
	uint32_t cap = (uint32_t)std::min<uint64_t>(3400, std::max<uint64_t>(small_max, max_rq + max_rq / 4 + 64));
	while (cap > 256 && join_lds(cap, false) > lds_max) cap -= 256;
	const size_t jlds = join_lds(cap, false);

	// ---- device buffers owned by the batch
	int rc = HJ_OK;
	auto fail = [&](int code) {
		batch_free(b);
		return code;
	};
	if ((rc = hj_reserve(ctx, b->part, part_n * W + 64))) return fail(rc);
	if ((rc = hj_reserve(ctx, b->hist, (hist_len + 1) * 8))) return fail(rc);
	if ((rc = hj_reserve(ctx, b->scan_tmp, ((hist_len + npairs) / SCAN_CH + 4) * 8 + 4096))) return fail(rc);
	if ((rc = hj_reserve(ctx, b->stats, sizeof(hj_stats)))) return fail(rc);
	if ((rc = hj_reserve(ctx, b->res, (size_t)njoins * 20 + 64))) return fail(rc);
	size_t off = 0;
	auto take = [&](size_t bytes) {
		size_t o = off;
		off += (bytes + 255) & ~(size_t)255;
		return o;
	};
	const size_t o_segs = take(nseg * sizeof(hj_seg)), o_cmap = take(nchunks * 4), o_pb = take((pb_total + 1) * 8);
	const size_t o_joins = take(njoins * sizeof(hj_bjoin)), o_pj = take((size_t)npairs * 4);
	const size_t o_pairs = take((size_t)npairs * sizeof(hj_bpair)), o_tiles = take(((size_t)npairs + 1) * 8);
	if ((rc = hj_reserve(ctx, b->meta, off))) return fail(rc);
	char *meta = (char *)b->meta.p;
	hj_seg *d_segs = (hj_seg *)(meta + o_segs);
	uint32_t *d_cmap = (uint32_t *)(meta + o_cmap);
	uint64_t *d_pb = (uint64_t *)(meta + o_pb), *d_tiles = (uint64_t *)(meta + o_tiles);
	hj_bjoin *d_joins = (hj_bjoin *)(meta + o_joins);
	uint32_t *d_pj = (uint32_t *)(meta + o_pj);
	hj_bpair *d_pairs = (hj_bpair *)(meta + o_pairs);
	b->d_res = (unsigned long long *)b->res.p;
	b->d_jflags = (unsigned int *)((char *)b->res.p + (size_t)njoins * 16);
	hj_stats *d_st = (hj_stats *)b->stats.p;
	auto up = [&](void *dst, const void *src, size_t bytes) -> int {
		if (!bytes) return HJ_OK;
		hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
		return e == hipSuccess ? HJ_OK : hj_fail(ctx, HJ_EHIP, "batch upload: %s", hipGetErrorString(e));
	};
	if ((rc = up(d_segs, segs.data(), nseg * sizeof(hj_seg))) || (rc = up(d_cmap, cmap.data(), nchunks * 4)) ||
	    (rc = up(d_joins, joins.data(), njoins * sizeof(hj_bjoin))) || (rc = up(d_pj, pair_join.data(), npairs * 4)))
		return fail(rc);
	b->info.scratch_bytes = b->part.cap + b->hist.cap + b->scan_tmp.cap + b->meta.cap + b->res.cap + b->stats.cap;

	// kernel attributes are set before capture (not stream work)
	const void *jk = join_b_kernel(wv);
	allow_lds(ctx, jk);
	if (jlds + static_lds(jk) > ctx->prop.sharedMemPerBlock) return fail(hj_fail(ctx, HJ_EINVAL, "batch LDS table too big"));
	if (nseg && slds + 1024 > ctx->prop.sharedMemPerBlock) return fail(hj_fail(ctx, HJ_EINVAL, "batch radix bits too big"));
	if (nseg) set_scatter_attr(ctx, wv);

	// ---- capture the batch as one graph
	const bool kt = ctx->kernel_times;
	ctx->kernel_times = false;
	if (hipEventCreate(&b->e0) != hipSuccess || hipEventCreate(&b->e1) != hipSuccess) {
		(void)hipGetLastError();
		ctx->kernel_times = kt;
		return fail(hj_fail(ctx, HJ_EHIP, "event create"));
	}
	hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
	if (e != hipSuccess) {
		ctx->kernel_times = kt;
		return fail(hj_fail(ctx, HJ_EHIP, "begin capture: %s", hipGetErrorString(e)));
	}
	hipLaunchKernelGGL(k_batch_reset, dim3(std::max<uint32_t>(1, std::min<uint32_t>(1024, (njoins + 255) / 256))),
	                   dim3(256), 0, ctx->stream, d_st, nchunks, b->d_res, njoins, b->d_jflags);
	if (nseg) {
		rc = launch_pass(ctx, wv, variant, (unsigned)nchunks, d_segs, d_cmap, ch, 0, maxbits, (uint64_t *)b->hist.p, hist_len,
		                 b->part.p, d_st, (uint64_t *)b->scan_tmp.p, true);
		if (!rc)
			hipLaunchKernelGGL(k_bounds_b, dim3(std::min<uint32_t>(nseg, 65535)), dim3(256), 0, ctx->stream, d_segs, nseg,
			                   (const uint64_t *)b->hist.p, d_pb);
	}
	if (!rc && npairs) {
		hipLaunchKernelGGL(k_plan_b, dim3(std::min<uint32_t>(65535, (npairs + 255) / 256)), dim3(256), 0, ctx->stream,
		                   d_joins, d_pj, npairs, d_segs, d_pb, (const uint4 *)b->part.p, d_pairs, d_tiles, d_st,
		                   nseg && wv == 1);
		rc = scan_u64(ctx, d_tiles, npairs, (uint64_t *)b->scan_tmp.p);
		const int per_cu = jlds <= 40 * 1024 ? 4 : (jlds <= 76 * 1024 ? 2 : 1);
		const unsigned jgrid = (unsigned)std::max<uint64_t>(1, (uint64_t)ctx->num_cu * per_cu);
		if (!rc) launch_join_b(ctx, wv, jgrid, jlds, d_pairs, d_tiles, npairs, d_joins, cap, d_st, b->d_res);
	}
	hipGraph_t g = nullptr;
	e = hipStreamEndCapture(ctx->stream, &g);
	ctx->kernel_times = kt;
	if (rc || e != hipSuccess) {
		if (g) (void)hipGraphDestroy(g);
		return fail(rc ? rc : hj_fail(ctx, HJ_EHIP, "end capture: %s", hipGetErrorString(e)));
	}
	b->graph = g;
	e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
	if (e != hipSuccess) return fail(hj_fail(ctx, HJ_EHIP, "graph instantiate: %s", hipGetErrorString(e)));
	*out = b;
	return HJ_OK;
}

extern "C" int hj_batch_run(hj_batch_t *b, hj_join_result_t *results, hj_batch_info_t *info) {
	if (!b) return HJ_EINVAL;
	hj_ctx_t *ctx = b->ctx;
	if (b->njoins && !results) return hj_fail(ctx, HJ_EINVAL, "null results");
	std::lock_guard<std::mutex> lk(ctx->mu);
	HJ_HIP(ctx, hipSetDevice(ctx->device));
	if (b->exact) {
		b->info.device_ms = b->info.fallback_ms = 0;
		b->info.algo_bytes = 0;
		b->info.n_fallback = b->njoins;
		b->info.n_ktimes = 0;
		for (uint32_t j = 0; j < b->njoins; j++) {
			hj_result_t r{};
			const bool kt = ctx->kernel_times;
			ctx->kernel_times = false;
			const int rc = hj_run_lp_exact(ctx, &b->R[j], &b->S[j], &r);
			ctx->kernel_times = kt;
			if (rc) return rc;
			results[j].matched_cnt = r.matched_cnt;
			results[j].check_sum = r.check_sum;
			b->info.device_ms += r.total_time * 1e3;
			b->info.algo_bytes += r.algo_bytes;
		}
		b->info.fallback_ms = b->info.device_ms;
		if (info) *info = b->info;
		return HJ_OK;
	}
	HJ_HIP(ctx, hipEventRecord(b->e0, ctx->stream));
	if (b->timed) {
		ctx->kernel_times = true;
		ctx->n_kt = 0;
		ctx->kt_open = false;
		b->launch();
		HJ_HIP(ctx, hipGetLastError());
	} else {
		HJ_HIP(ctx, hipGraphLaunch(b->exec, ctx->stream));
	}
	HJ_HIP(ctx, hipEventRecord(b->e1, ctx->stream));
	std::vector<unsigned char> host((size_t)b->njoins * 20 + 8);
	if (b->njoins)
		HJ_HIP(ctx, hipMemcpyAsync(host.data(), b->res.p, (size_t)b->njoins * 20, hipMemcpyDeviceToHost, ctx->stream));
	hj_stats hs{};
	HJ_HIP(ctx, hipMemcpyAsync(&hs, b->stats.p, sizeof(hj_stats), hipMemcpyDeviceToHost, ctx->stream));
	HJ_HIP(ctx, hipStreamSynchronize(ctx->stream));
	if (b->gathered) {
		// pass A reads 16 B and writes each tile at its layout; the join reads that
		const uint64_t N = b->part_tuples, nt = hs.narrow_tuples, nd = hs.dense_tuples;
		b->info.algo_bytes = 16 * N + 2 * (4 * nd + 8 * (nt - nd) + 16 * (N - nt));
	} else {
		// the scatter reads W and writes W (8 B for narrow 16-byte tuples), the
		// join reads that; unpartitioned joins are read once (the histogram
		// pass is not counted)
		const uint64_t out = (b->W == 16 && !(hs.flags & HJ_DF_WIDE)) ? 8 : b->W;
		b->info.algo_bytes = b->part_tuples * (b->W + 2 * out) + b->unpart_tuples * b->W;
	}
	float ms = 0.f;
	HJ_HIP(ctx, hipEventElapsedTime(&ms, b->e0, b->e1));
	b->info.device_ms = ms;
	b->info.n_ktimes = 0;
	if (b->timed) {
		b->info.n_ktimes = (uint32_t)ctx->n_kt;
		for (int i = 0; i < ctx->n_kt; i++) {
			b->info.ktime_id[i] = ctx->kt_id[i];
			b->info.ktime_ms[i] = hj_kt_ms(ctx, i);
		}
		ctx->kernel_times = false;
	}
	b->info.fallback_ms = 0;
	b->info.n_fallback = 0;
	const unsigned long long *res = (const unsigned long long *)host.data();
	const unsigned int *fl = (const unsigned int *)(host.data() + (size_t)b->njoins * 16);
	for (uint32_t j = 0; j < b->njoins; j++) {
		results[j].matched_cnt = res[2 * j];
		results[j].check_sum = res[2 * j + 1];
		if (b->lp && (fl[j] & (HJ_DF_DUPKEY | HJ_DF_OVERSIZE))) {
			// exact first-match path for this join, outside the graph
			hj_result_t r{};
			const bool kt = ctx->kernel_times;
			ctx->kernel_times = false;
			int rc = hj_run_lp_exact(ctx, &b->R[j], &b->S[j], &r);
			ctx->kernel_times = kt;
			if (rc) return rc;
			results[j].matched_cnt = r.matched_cnt;
			results[j].check_sum = r.check_sum;
			b->info.n_fallback++;
			b->info.fallback_ms += r.total_time * 1e3;
		}
	}
	if (info) *info = b->info;
	return HJ_OK;
}

extern "C" int hj_batch_destroy(hj_batch_t *b) {
	if (!b) return HJ_EINVAL;
	hj_ctx_t *ctx = b->ctx;
	std::lock_guard<std::mutex> lk(ctx->mu);
	(void)hipSetDevice(ctx->device);
	(void)hipStreamSynchronize(ctx->stream);
	batch_free(b);
	return HJ_OK;
}

// ---------------------------------------------------------------- autotuner
extern "C" int hj_autotune(hj_ctx_t *ctx, uint32_t width, uint32_t s_ratio, uint32_t min_log2, uint32_t max_log2) {
	if (!ctx || s_ratio == 0 || min_log2 > max_log2 || max_log2 > 28) return hj_fail(ctx, HJ_EINVAL, "bad tuning range");
	if (width != 16 && width != 32 && width != 64 && width != 128 && width != 256)
		return hj_fail(ctx, HJ_EINVAL, "tuple size %u", width);
	std::unique_lock<std::mutex> lk(ctx->mu);
	HJ_HIP(ctx, hipSetDevice(ctx->device));
	const uint64_t nr_max = 1ull << max_log2, ns_max = nr_max * s_ratio;
	void *r = nullptr, *s = nullptr;
	if (hipMalloc(&r, nr_max * width) != hipSuccess || hipMalloc(&s, ns_max * width) != hipSuccess) {
		(void)hipGetLastError();
		if (r) (void)hipFree(r);
		return hj_fail(ctx, HJ_ENOMEM, "autotune buffers");
	}
	int rc = HJ_OK;
	const bool kt = ctx->kernel_times;
	ctx->kernel_times = false;
	for (uint32_t c = min_log2; c <= max_log2 && rc == HJ_OK; c++) {
		hj_gen_spec_t spec{};
		spec.workload = HJ_WL_PKFK_UNIFORM;
		spec.width = width;
		spec.seed = 777 + c;
		spec.r_num = 1ull << c;
		spec.s_num = (1ull << c) * s_ratio;
		if ((rc = hj_launch_generate(ctx, &spec, r, s))) break;
		hj_relation_t R{r, spec.r_num, width, HJ_LOC_DEVICE}, S{s, spec.s_num, width, HJ_LOC_DEVICE};
		double best = 1e30, best1 = 1e30;
		uint32_t bb = 0, bp = 0, bs = 0, bb1 = 0;
		// candidates: partitions of 512..8192 build tuples, one or two passes,
		// and for two passes the first-pass width around the even split
		// (asymmetric bits, phj_asymm_radix.h:10-22)
		for (int b = (int)c - 13; b <= (int)c - 9; b++) {
			if (b < 0) continue;
			for (uint32_t q = 0; q < 4; q++) {
				const uint32_t p = q == 0 ? (b == 0 ? 0u : 1u) : 2u;
				if (p == 2 && b < 6) break;  // two passes below 6 bits never win
				if (b > 12 && p == 1) continue;
				// q = 1, 2, 3: first pass takes the even split -1, +0, +1
				const uint32_t s1 = p == 2 ? (uint32_t)(b + 1) / 2 + q - 2 : 0;
				if (p == 2 && (s1 < 1 || s1 >= (uint32_t)b || s1 > 12 || b - s1 > 12)) continue;
				hj_params_t prm{};
				prm.radix_bits = (uint32_t)b;
				prm.passes = p;
				prm.radix_bits_pass1 = s1;
				if (b == 0) prm.passes = 0;
				double t = 1e30;
				for (int rep = 0; rep < 3; rep++) {
					hj_result_t out{};
					rc = hj_run_radix(ctx, &R, &S, &prm, false, &out);
					if (rc) break;
					t = std::min(t, out.total_time);
				}
				if (rc == HJ_EINVAL) {  // configuration not supported here: skip it
					rc = HJ_OK;
					continue;
				}
				if (rc) break;
				if (t < best) {
					best = t;
					bb = (uint32_t)b;
					bp = b ? p : 0;
					bs = s1;
				}
				if (p <= 1 && t < best1) {  // the one-pass choice of batches and serving
					best1 = t;
					bb1 = (uint32_t)b;
				}
			}
			if (rc) break;
		}
		ctx->tune_bits[c] = (uint8_t)bb;
		ctx->tune_passes[c] = (uint8_t)bp;
		ctx->tune_pass1[c] = (uint8_t)bs;
		ctx->tune_bits1[c] = (uint8_t)bb1;
		ctx->tune_var1[c] = 0;
		ctx->tuned = true;
		// The one-pass scatter's tile shape and write-combining line for this
		// class, timed through a batch of this one join (the path that uses
		// them: batches and serving partition in one pass).  hj_batch_* take
		// the context's lock themselves.
		if (rc == HJ_OK && bb1) {
			static const uint32_t k_vars[] = {0, 1, 2, 3, 4, 5, 3 << 4, 4 << 4, 5 << 4};
			double bestv = 1e30;
			lk.unlock();
			for (uint32_t v : k_vars) {
				hj_params_t prm{};
				prm.part_variant = v;
				prm.flags = HJ_F_BATCH_SCATTER;  // the shapes of the scatter path
				hj_batch_t *bt = nullptr;
				if (hj_batch_create(ctx, "phj_radix_bc", 1, &R, &S, &prm, &bt) != HJ_OK) continue;
				double t = 1e30;
				for (int rep = 0; rep < 3; rep++) {
					hj_join_result_t jr{};
					hj_batch_info_t bi{};
					if (hj_batch_run(bt, &jr, &bi) != HJ_OK) break;
					t = std::min(t, bi.device_ms);
				}
				hj_batch_destroy(bt);
				if (t < bestv) {
					bestv = t;
					ctx->tune_var1[c] = (uint8_t)v;
				}
			}
			lk.lock();
		}
		// Gathered batches (16-byte bc): the radix bits of this class, the
		// 6144-tuple rule and one bit either side, timed through a gathered
		// batch of this one join
		if (rc == HJ_OK && width == 16 && spec.r_num <= (1ull << 22)) {
			const uint32_t b0 = gbatch_bits(spec.r_num);
			double bestg = 1e30;
			uint32_t bestb = b0;
			lk.unlock();
			for (int db = -1; db <= 1; db++) {
				const int b = (int)b0 + db;
				if (b < 0 || b > 10) continue;
				ctx->tune_gbits[c] = (uint8_t)b;
				hj_batch_t *bt = nullptr;
				if (hj_batch_create(ctx, "phj_radix_bc", 1, &R, &S, nullptr, &bt) != HJ_OK) continue;
				double t = 1e30;
				for (int rep = 0; rep < 3; rep++) {
					hj_join_result_t jr{};
					hj_batch_info_t bi{};
					if (hj_batch_run(bt, &jr, &bi) != HJ_OK) break;
					t = std::min(t, bi.device_ms);
				}
				hj_batch_destroy(bt);
				if (t < bestg) {
					bestg = t;
					bestb = (uint32_t)b;
				}
			}
			ctx->tune_gbits[c] = (uint8_t)bestb;
			lk.lock();
		}
	}
	ctx->kernel_times = kt;
	(void)hipStreamSynchronize(ctx->stream);
	(void)hipFree(r);
	(void)hipFree(s);
	if (rc == HJ_OK) ctx->tuned = true;
	return rc;
}

// Skew-aware tuning: the same (bits, passes, first-pass bits) candidates as
// hj_autotune, timed on probe sides whose keys are Zipf-distributed over the
// build keys (a few heavy partitions), then the probe-range scale for the
// best of them: a smaller scale spreads a heavy partition over more work
// items, a larger one rebuilds each partition's table less often.  The
// reference's counterpart is its SKEW_HANDLING build (phj.c:1374-1669:
// re-partitioning large partitions and splitting large probe sides), a
// compile-time choice; here joins opt in with HJ_F_SKEW_HANDLING.
extern "C" int hj_autotune_zipf(hj_ctx_t *ctx, uint32_t width, uint32_t s_ratio, uint32_t min_log2, uint32_t max_log2,
                                double theta) {
	if (!ctx || s_ratio == 0 || min_log2 > max_log2 || max_log2 > 28 || !(theta >= 0.0 && theta <= 4.0))
		return hj_fail(ctx, HJ_EINVAL, "bad tuning range");
	if (width != 16 && width != 32 && width != 64 && width != 128 && width != 256)
		return hj_fail(ctx, HJ_EINVAL, "tuple size %u", width);
	std::lock_guard<std::mutex> lk(ctx->mu);
	HJ_HIP(ctx, hipSetDevice(ctx->device));
	const uint64_t nr_max = 1ull << max_log2, ns_max = nr_max * s_ratio;
	void *r = nullptr, *s = nullptr;
	if (hipMalloc(&r, nr_max * width) != hipSuccess || hipMalloc(&s, ns_max * width) != hipSuccess) {
		(void)hipGetLastError();
		if (r) (void)hipFree(r);
		return hj_fail(ctx, HJ_ENOMEM, "autotune buffers");
	}
	int rc = HJ_OK;
	const bool kt = ctx->kernel_times;
	ctx->kernel_times = false;
	auto time_of = [&](const hj_relation_t &R, const hj_relation_t &S, const hj_params_t &prm, double &t) {
		t = 1e30;
		for (int rep = 0; rep < 3; rep++) {
			hj_result_t out{};
			const int e = hj_run_radix(ctx, &R, &S, &prm, false, &out);
			if (e) return e;
			t = std::min(t, out.total_time);
		}
		return HJ_OK;
	};
	for (uint32_t c = min_log2; c <= max_log2 && rc == HJ_OK; c++) {
		hj_gen_spec_t spec{};
		spec.workload = HJ_WL_ZIPF;
		spec.width = width;
		spec.seed = 991 + c;
		spec.r_num = 1ull << c;
		spec.s_num = (1ull << c) * s_ratio;
		spec.theta = theta;
		if ((rc = hj_launch_generate(ctx, &spec, r, s))) break;
		hj_relation_t R{r, spec.r_num, width, HJ_LOC_DEVICE}, S{s, spec.s_num, width, HJ_LOC_DEVICE};
		double best = 1e30;
		hj_params_t bestp{};
		bool found = false;
		for (int b = (int)c - 13; b <= (int)c - 9; b++) {
			if (b < 0) continue;
			for (uint32_t q = 0; q < 4; q++) {
				const uint32_t p = q == 0 ? (b == 0 ? 0u : 1u) : 2u;
				if (p == 2 && b < 6) break;
				if (b > 12 && p == 1) continue;
				const uint32_t s1 = p == 2 ? (uint32_t)(b + 1) / 2 + q - 2 : 0;
				if (p == 2 && (s1 < 1 || s1 >= (uint32_t)b || s1 > 12 || b - s1 > 12)) continue;
				hj_params_t prm{};
				prm.radix_bits = (uint32_t)b;
				prm.passes = b ? p : 0;
				prm.radix_bits_pass1 = s1;
				double t;
				rc = time_of(R, S, prm, t);
				if (rc == HJ_EINVAL) {  // not supported here: skip it
					rc = HJ_OK;
					continue;
				}
				if (rc) break;
				if (t < best) {
					best = t;
					bestp = prm;
					found = true;
				}
			}
			if (rc) break;
		}
		if (rc || !found) continue;
		uint32_t bests = 0;
		for (uint32_t sc = 1; sc <= 3 && bestp.radix_bits; sc++) {
			hj_params_t prm = bestp;
			prm.part_variant = sc << 20;
			double t;
			rc = time_of(R, S, prm, t);
			if (rc) break;
			if (t < best) {
				best = t;
				bests = sc;
			}
		}
		if (rc) break;
		ctx->tune_bits_z[c] = (uint8_t)bestp.radix_bits;
		ctx->tune_passes_z[c] = (uint8_t)bestp.passes;
		ctx->tune_pass1_z[c] = (uint8_t)bestp.radix_bits_pass1;
		ctx->tune_pscale_z[c] = (uint8_t)bests;
		ctx->tuned_z = true;
	}
	ctx->kernel_times = kt;
	(void)hipStreamSynchronize(ctx->stream);
	(void)hipFree(r);
	(void)hipFree(s);
	return rc;
}

extern "C" int hj_tuned_skew_config(hj_ctx_t *ctx, uint64_t r_num, uint32_t *radix_bits, uint32_t *passes,
                                    uint32_t *radix_bits_pass1, uint32_t *probe_scale) {
	if (!ctx || !radix_bits || !passes || !radix_bits_pass1 || !probe_scale) return HJ_EINVAL;
	hj_params_t prm{};
	prm.flags = HJ_F_SKEW_HANDLING;
	radix_choose(ctx, r_num, &prm, radix_bits, passes, radix_bits_pass1, probe_scale);
	return HJ_OK;
}

extern "C" int hj_tuned_batch_bits(hj_ctx_t *ctx, uint64_t r_num, uint32_t *radix_bits) {
	if (!ctx || !radix_bits) return HJ_EINVAL;
	*radix_bits = gbatch_bits(ctx, r_num);
	return HJ_OK;
}

extern "C" int hj_tuned_config(hj_ctx_t *ctx, uint64_t r_num, uint32_t *radix_bits, uint32_t *passes) {
	if (!ctx || !radix_bits || !passes) return HJ_EINVAL;
	hj_radix_config(ctx, r_num, radix_bits, passes);
	return HJ_OK;
}

extern "C" int hj_tuned_pass1_bits(hj_ctx_t *ctx, uint64_t r_num, uint32_t *radix_bits_pass1) {
	if (!ctx || !radix_bits_pass1) return HJ_EINVAL;
	uint32_t b, p;
	hj_radix_config(ctx, r_num, &b, &p, radix_bits_pass1);
	return HJ_OK;
}

// ======================================================================
// Serving: continuous batching over per-slot streams and HBM arenas.
// ======================================================================
#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <unordered_map>

typedef std::chrono::steady_clock hj_clock;

struct hj_build_res {
	hj_relation_t R{};
	uint32_t bits = 0;
	void *part = nullptr;    // partitioned copy of R (bits > 0)
	uint64_t *pb = nullptr;  // 2^bits + 1 partition bounds
	size_t bytes = 0;
	// gathered servers: R's tiles sorted by the digit, run rows, descriptors
	bool gathered = false;
	hj_gres g;
	void *gbuf = nullptr;
};

struct hj_sreq {
	int32_t build = -1;
	hj_relation_t R{}, S{};
	uint64_t footprint = 0;
	bool gathered = false;  // runs in a gathered batch
	hj_clock::time_point t_arr, t_done;
	int state = 0;  // 0 queued, 1 running, 2 done
	int err = HJ_OK; // status reported by poll for a request that could not run
	hj_join_result_t res{};
};

struct hj_slot {
	hipStream_t stream = nullptr;
	hipEvent_t done = nullptr;
	char *arena = nullptr;
	size_t arena_bytes = 0;
	unsigned char *hbuf = nullptr;  // pinned: plan upload, then results
	size_t hbuf_bytes = 0;
	size_t res_off = 0;
	std::vector<uint64_t> reqs;
	uint64_t admitted = 0;
	bool busy = false;
	bool big = false;  // the batch in flight holds a large request
};

struct hj_solo_lane {
	hj_ctx_t *ctx = nullptr;  // own stream and scratch
	std::thread th;
};

struct hj_server {
	hj_ctx_t *ctx = nullptr;
	bool lp = false;
	std::string algo;
	std::vector<hj_solo_lane> solo;
	std::deque<uint64_t> solo_queue;
	uint64_t solo_min = 1ull << 16;
	uint64_t solo_budget = 0;  // scratch one solo lane may use
	hj_serve_config_t cfg{};
	uint32_t W = 0;
	std::vector<hj_slot> slots;
	std::vector<hj_build_res> builds;
	std::mutex mu;
	std::condition_variable cv;
	std::deque<uint64_t> queue;
	std::unordered_map<uint64_t, hj_sreq> reqs;
	uint64_t next_id = 1;
	size_t solo_busy = 0;  // lanes running a request
	bool stop = false;
	std::thread th;
	hj_serve_stats_t st{};
	uint64_t admitted_now = 0;
};

// tuples per partition chunk when serving: a multiple of every scatter tile
static const uint64_t kServeChunk = 16384;

static size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

// Gathered serving: 16-byte servers run their batches gathered (first-
// pass tiles, gathered joins), like gathered hj_batch_* batches.
static bool serve_gathered(const hj_server *s) { return s->W == 16 && !(s->cfg.flags & HJ_SERVE_SCATTER); }

// Largest build side a gathered server batches (the gathered limit).
static uint64_t serve_gmax(const hj_server *) { return 1ull << 22; }
// Above 2^20 build tuples a partition's run in an 8192-tuple tile gets
// shorter than 32 tuples and a lane's two-pass engine is faster alone, so
// such a request goes to a solo lane while one is idle (low load: the
// latency of the two-pass engine) and into a gathered batch otherwise
// (saturation: the batches' throughput).
static const uint64_t kServeSoloGathered = 1ull << 20;
// A request of at least this many tuples (|R| + |S|, about half a
// millisecond of device time) runs alone while fewer than kServeHoldQueue
// requests wait (see serve_loop).
static const uint64_t kServeBigTuples = 1ull << 26;
static const size_t kServeHoldQueue = 16;

// Whether request r can join a gathered batch: its build side is a resident
// build kept in gathered layout, or holds at most serve_gmax tuples, and its
// probe side fewer than 2^30.
static bool serve_req_gatherable(const hj_server *s, const hj_sreq &r) {
	if (!serve_gathered(s) || r.S.tuple_num >= (1ull << 30)) return false;
	return r.build >= 0 ? s->builds[r.build].gathered : r.R.tuple_num <= serve_gmax(s);
}

// Scratch a gathered request adds to a batch: its plan alone (the batch's
// regions are aligned once, and extra probe-range items when the batch is
// small are within the batch overhead).
static uint64_t serve_gfootprint(const hj_server *s, const hj_sreq &r) {
	hj_gplan P;
	gplan_add(s->ctx, P, r.R, r.S, r.build >= 0 ? &s->builds[r.build].g : nullptr);
	gplan_finish(P, 0);
	return P.bytes + 20;
}

// Exact scratch a request adds to a batch (its share of every arena region).
struct serve_req_plan {
	uint32_t bits;
	bool r_here;  // R partitioned in this batch
	uint64_t part_tuples, nchunks, hist_entries, pb_entries, npairs;
	uint32_t nseg;
};

static serve_req_plan serve_plan_req(hj_server *s, const hj_sreq &r) {
	serve_req_plan p{};
	const uint64_t nr = r.R.tuple_num, ns = r.S.tuple_num;
	if (r.build >= 0) {
		p.bits = s->builds[r.build].bits;
		p.r_here = false;
	} else {
		p.bits = nr > 2048 ? hj_radix_bits_onepass(s->ctx, nr) : 0;
		p.r_here = p.bits > 0;
	}
	const uint64_t F = 1ull << p.bits;
	p.npairs = F;
	if (p.bits) {
		const uint64_t nsides[2] = {p.r_here ? nr : UINT64_MAX, ns};
		for (uint64_t len : nsides) {
			if (len == UINT64_MAX) continue;
			const uint64_t c = (len + kServeChunk - 1) / kServeChunk;
			p.part_tuples += len;
			p.nchunks += c;
			p.hist_entries += c * F;
			p.pb_entries += F + 1;
			p.nseg++;
		}
	}
	return p;
}

static uint64_t serve_footprint(const serve_req_plan &p, uint32_t W) {
	return p.part_tuples * W + p.hist_entries * 8 + p.nchunks * 4 + p.pb_entries * 8 + p.nseg * sizeof(hj_seg) +
	       sizeof(hj_bjoin) + p.npairs * (sizeof(hj_bpair) + 4 + 8) + 20 + (p.hist_entries + p.npairs) / 512 + 64;
}

static const uint64_t kServeBatchOverhead = 1u << 20;  // alignment, scan partials, stats

static int serve_grow_pinned(hj_server *s, hj_slot &sl, size_t bytes) {
	if (sl.hbuf_bytes >= bytes) return HJ_OK;
	if (sl.hbuf) (void)hipHostFree(sl.hbuf);
	sl.hbuf = nullptr;
	sl.hbuf_bytes = 0;
	size_t want = std::max<size_t>(bytes, 1u << 20);
	if (hipHostMalloc((void **)&sl.hbuf, want, hipHostMallocDefault) != hipSuccess) {
		(void)hipGetLastError();
		return hj_fail(s->ctx, HJ_ENOMEM, "pinned staging buffer");
	}
	sl.hbuf_bytes = want;
	return HJ_OK;
}

// Plan and enqueue one batch of queued requests on a free slot.  Called with
// s->mu held; takes ctx->mu for the launches.
static int serve_launch(hj_server *s, hj_slot &sl, const std::vector<uint64_t> &ids) {
	hj_ctx_t *ctx = s->ctx;
	const uint32_t n = (uint32_t)ids.size(), W = s->W, wv = W / 16;
	std::vector<hj_bjoin> joins(n);
	std::vector<hj_seg> segs;
	std::vector<uint32_t> cmap, pj;
	uint64_t part_n = 0, hist_len = 0, max_rq = 1;
	uint32_t maxbits = 0, pb_total = 0;
	for (uint32_t j = 0; j < n; j++) {
		const hj_sreq &r = s->reqs[ids[j]];
		const serve_req_plan p = serve_plan_req(s, r);
		hj_bjoin &J = joins[j];
		J.R = (const uint4 *)r.R.tuples;
		J.S = (const uint4 *)r.S.tuples;
		J.nr = r.R.tuple_num;
		J.ns = r.S.tuple_num;
		J.bits = p.bits;
		J.seg = (uint32_t)segs.size();
		J.pair_base = (uint32_t)pj.size();
		J.rpb = nullptr;
		J.rpart = nullptr;
		if (r.build >= 0 && p.bits) {
			J.rpb = s->builds[r.build].pb;
			J.rpart = (const uint4 *)s->builds[r.build].part;
		}
		J.s_tile = std::max<uint64_t>(8192, 2 * ((J.ns >> p.bits) + 1));
		max_rq = std::max<uint64_t>(max_rq, (J.nr >> p.bits) + 1);
		for (uint64_t q = 0; q < (1ull << p.bits); q++) pj.push_back(j);
		if (p.bits) {
			maxbits = std::max(maxbits, p.bits);
			for (int side = p.r_here ? 0 : 1; side < 2; side++) {
				hj_seg g{};
				g.in = side ? J.S : J.R;
				g.len = side ? J.ns : J.nr;
				g.out_start = part_n;
				g.bits = p.bits;
				g.pb_base = pb_total;
				pb_total += (1u << p.bits) + 1;
				g.chunk_base = (uint32_t)cmap.size();
				g.nchunks = (uint32_t)((g.len + kServeChunk - 1) / kServeChunk);
				g.hist_base = hist_len;
				hist_len += (uint64_t)g.nchunks << p.bits;
				for (uint32_t c = 0; c < g.nchunks; c++) cmap.push_back((uint32_t)segs.size());
				part_n += g.len;
				segs.push_back(g);
			}
		}
	}
	const uint32_t nseg = (uint32_t)segs.size(), npairs = (uint32_t)pj.size();
	const uint64_t nchunks = cmap.size();
	// arena layout
	size_t off = 0;
	auto take = [&](size_t b) {
		size_t o = off;
		off += al256(b);
		return o;
	};
	const size_t o_part = take(part_n * W + 64), o_hist = take((hist_len + 1) * 8);
	const size_t o_scan = take(((hist_len + npairs) / SCAN_CH + 4) * 8 + 4096), o_st = take(sizeof(hj_stats));
	const size_t o_res = take((size_t)n * 20 + 64), o_pb = take(((size_t)pb_total + 1) * 8);
	const size_t o_pairs = take((size_t)npairs * sizeof(hj_bpair)), o_tiles = take(((size_t)npairs + 1) * 8);
	// uploaded together: segs | cmap | joins | pair->join
	const size_t up0 = off;
	const size_t o_segs = take(nseg * sizeof(hj_seg)), o_cmap = take(nchunks * 4);
	const size_t o_joins = take((size_t)n * sizeof(hj_bjoin)), o_pj = take((size_t)npairs * 4);
	const size_t up_bytes = off - up0;
	if (off > sl.arena_bytes) return hj_fail(ctx, HJ_ENOMEM, "batch does not fit its arena (%zu > %zu)", off, sl.arena_bytes);
	int rc = serve_grow_pinned(s, sl, up_bytes + (size_t)n * 20 + 4096);
	if (rc) return rc;
	memcpy(sl.hbuf + (o_segs - up0), segs.data(), nseg * sizeof(hj_seg));
	memcpy(sl.hbuf + (o_cmap - up0), cmap.data(), nchunks * 4);
	memcpy(sl.hbuf + (o_joins - up0), joins.data(), (size_t)n * sizeof(hj_bjoin));
	memcpy(sl.hbuf + (o_pj - up0), pj.data(), (size_t)npairs * 4);
	sl.res_off = al256(up_bytes);
	char *A = sl.arena;
	hj_seg *d_segs = (hj_seg *)(A + o_segs);
	uint32_t *d_cmap = (uint32_t *)(A + o_cmap), *d_pj = (uint32_t *)(A + o_pj);
	hj_bjoin *d_joins = (hj_bjoin *)(A + o_joins);
	hj_bpair *d_pairs = (hj_bpair *)(A + o_pairs);
	uint64_t *d_tiles = (uint64_t *)(A + o_tiles), *d_pb = (uint64_t *)(A + o_pb), *d_hist = (uint64_t *)(A + o_hist);
	hj_stats *d_st = (hj_stats *)(A + o_st);
	unsigned long long *d_res = (unsigned long long *)(A + o_res);
	unsigned int *d_fl = (unsigned int *)(A + o_res + (size_t)n * 16);
	const size_t lds_max =
	    std::min<size_t>(ctx->prop.sharedMemPerBlock, 160 * 1024) - static_lds(join_b_kernel(wv)) - 256;
	// as in batches: at most 3400 tuples, two workgroups per CU
	uint32_t cap = (uint32_t)std::min<uint64_t>(3400, std::max<uint64_t>(2048, max_rq + max_rq / 4 + 64));
	while (cap > 256 && join_lds(cap, false) > lds_max) cap -= 256;
	const size_t jlds = join_lds(cap, false);
	if (jlds > lds_max) return hj_fail(ctx, HJ_EINVAL, "serve join table does not fit LDS");

	std::lock_guard<std::mutex> lk(ctx->mu);
	hipStream_t saved = ctx->stream;
	const bool kt = ctx->kernel_times;
	ctx->kernel_times = false;
	ctx->stream = sl.stream;
	hipError_t e = hipMemcpyAsync(A + up0, sl.hbuf, up_bytes, hipMemcpyHostToDevice, sl.stream);
	if (e == hipSuccess) {
		if (nseg) set_scatter_attr(ctx, wv);
		allow_lds(ctx, join_b_kernel(wv));
		hipLaunchKernelGGL(k_batch_reset, dim3(std::max<uint32_t>(1, std::min<uint32_t>(1024, (n + 255) / 256))),
		                   dim3(256), 0, sl.stream, d_st, nchunks, d_res, n, d_fl);
		// an lp server's requests that do not run gathered are joined by the
		// exact first-match path when the batch finishes: no device join here
		if (nseg && !s->lp) {
			rc = launch_pass(ctx, wv, 0, (unsigned)nchunks, d_segs, d_cmap, kServeChunk, 0, maxbits, d_hist, hist_len,
			                 A + o_part, d_st, (uint64_t *)(A + o_scan), true);
			if (!rc)
				hipLaunchKernelGGL(k_bounds_b, dim3(std::min<uint32_t>(nseg, 65535)), dim3(256), 0, sl.stream, d_segs,
				                   nseg, (const uint64_t *)d_hist, d_pb);
		}
		if (!rc && !s->lp) {
			hipLaunchKernelGGL(k_plan_b, dim3(std::min<uint32_t>(65535, (npairs + 255) / 256)), dim3(256), 0, sl.stream,
			                   d_joins, d_pj, npairs, d_segs, d_pb, (const uint4 *)(A + o_part), d_pairs, d_tiles, d_st,
			                   nseg && wv == 1);
			rc = scan_u64(ctx, d_tiles, npairs, (uint64_t *)(A + o_scan));
			if (!rc) {
				const int per_cu = jlds <= 40 * 1024 ? 4 : (jlds <= 76 * 1024 ? 2 : 1);
				launch_join_b(ctx, wv, (unsigned)ctx->num_cu * per_cu, jlds, d_pairs, d_tiles, npairs, d_joins, cap, d_st,
				              d_res);
			}
		}
		e = hipGetLastError();
		if (e == hipSuccess) e = hipMemcpyAsync(sl.hbuf + sl.res_off, d_res, (size_t)n * 20, hipMemcpyDeviceToHost, sl.stream);
		if (e == hipSuccess) e = hipEventRecord(sl.done, sl.stream);
	}
	ctx->stream = saved;
	ctx->kernel_times = kt;
	if (rc) return rc;
	if (e != hipSuccess) return hj_fail(ctx, HJ_EHIP, "serve launch: %s", hipGetErrorString(e));
	sl.reqs = ids;
	sl.busy = true;
	return HJ_OK;
}

// Plan and enqueue one gathered batch (every request gatherable) on a free
// slot: the plan's region, the stats block and the results in the slot's
// arena, the plan uploaded from the slot's pinned buffer.  Called with s->mu
// held; takes ctx->mu for the launches.
static int serve_launch_g(hj_server *s, hj_slot &sl, const std::vector<uint64_t> &ids) {
	hj_ctx_t *ctx = s->ctx;
	const uint32_t n = (uint32_t)ids.size();
	hj_gplan P;
	for (uint64_t id : ids) {
		const hj_sreq &r = s->reqs[id];
		gplan_add(ctx, P, r.R, r.S, r.build >= 0 ? &s->builds[r.build].g : nullptr);
	}
	gplan_finish(P, (uint32_t)ctx->num_cu);
	const size_t o_st = al256(P.bytes), o_res = o_st + al256(sizeof(hj_stats)), total = o_res + al256((size_t)n * 20 + 64);
	if (total > sl.arena_bytes) return hj_fail(ctx, HJ_ENOMEM, "batch does not fit its arena (%zu > %zu)", total, sl.arena_bytes);
	const size_t up_bytes = P.bytes - P.o_up;
	int rc = serve_grow_pinned(s, sl, al256(up_bytes) + (size_t)n * 20 + 4096);
	if (rc) return rc;
	char *A = sl.arena;
	gplan_place(P, A, sl.hbuf);
	sl.res_off = al256(up_bytes);
	hj_stats *d_st = (hj_stats *)(A + o_st);
	unsigned long long *d_res = (unsigned long long *)(A + o_res);
	unsigned int *d_fl = (unsigned int *)(A + o_res + (size_t)n * 16);
	std::lock_guard<std::mutex> lk(ctx->mu);
	hj_glaunch L;
	if ((rc = gplan_prepare(ctx, P, A, L, s->lp))) return rc;
	const bool kt = ctx->kernel_times;
	ctx->kernel_times = false;
	hipError_t e = hipMemcpyAsync(A + P.o_up, sl.hbuf, up_bytes, hipMemcpyHostToDevice, sl.stream);
	if (e == hipSuccess) {
		gplan_launch(ctx, sl.stream, L, d_st, d_res, d_fl);
		e = hipGetLastError();
	}
	if (e == hipSuccess) e = hipMemcpyAsync(sl.hbuf + sl.res_off, d_res, (size_t)n * 20, hipMemcpyDeviceToHost, sl.stream);
	if (e == hipSuccess) e = hipEventRecord(sl.done, sl.stream);
	ctx->kernel_times = kt;
	if (e != hipSuccess) return hj_fail(ctx, HJ_EHIP, "serve launch: %s", hipGetErrorString(e));
	sl.reqs = ids;
	sl.busy = true;
	return HJ_OK;
}

static void serve_finish(hj_server *s, hj_slot &sl) {
	const uint32_t n = (uint32_t)sl.reqs.size();
	const unsigned long long *res = (const unsigned long long *)(sl.hbuf + sl.res_off);
	const unsigned int *fl = (const unsigned int *)(sl.hbuf + sl.res_off + (size_t)n * 16);
	for (uint32_t j = 0; j < n; j++) {
		hj_sreq &r = s->reqs[sl.reqs[j]];
		r.res.matched_cnt = res[2 * j];
		r.res.check_sum = res[2 * j + 1];
		if (s->lp && (!r.gathered || (fl[j] & (HJ_DF_DUPKEY | HJ_DF_OVERSIZE)))) {
			// exact first-match path for this request (every request of an lp
			// server that did not run gathered; a gathered one that flagged a
			// duplicate build key or a multi-fill partition)
			std::lock_guard<std::mutex> lk(s->ctx->mu);
			hj_result_t out{};
			const hj_relation_t &R = r.build >= 0 ? s->builds[r.build].R : r.R;
			r.err = hj_run_lp_exact(s->ctx, &R, &r.S, &out);
			r.res.matched_cnt = out.matched_cnt;
			r.res.check_sum = out.check_sum;
		}
		r.t_done = hj_clock::now();
		r.state = 2;
		s->st.completed++;
		s->admitted_now -= r.footprint;
	}
	sl.reqs.clear();
	sl.busy = false;
}

static void serve_loop(hj_server *s) {
	(void)hipSetDevice(s->ctx->device);
	std::unique_lock<std::mutex> lk(s->mu);
	hj_clock::time_point last_active = hj_clock::now();
	for (;;) {
		bool progress = false;
		uint32_t busy = 0;
		for (hj_slot &sl : s->slots) {
			if (!sl.busy) continue;
			const hipError_t q = hipEventQuery(sl.done);
			if (q == hipSuccess) {
				serve_finish(s, sl);
				progress = true;
			} else {
				if (q != hipErrorNotReady) (void)hipGetLastError();
				busy++;
			}
		}
		// Light load: while a batch holding a large request runs and few
		// requests wait, no other batch starts, so the large one has the
		// device to itself (as one join at a time would) instead of sharing it
		// with the small ones arriving meanwhile; a longer queue brings the
		// overlap back for throughput.
		bool big_running = false;
		for (const hj_slot &sl : s->slots) big_running |= sl.busy && sl.big;
		const bool hold = big_running && s->queue.size() < kServeHoldQueue;
		for (hj_slot &sl : s->slots) {
			if (sl.busy || s->queue.empty() || hold) continue;
			std::vector<uint64_t> ids;
			uint64_t fp = kServeBatchOverhead;
			uint64_t hist_entries = 0;
			// a batch is all gathered or all scattered, like the queue's front
			const bool g = s->reqs[s->queue.front()].gathered;
			while (!s->queue.empty() && ids.size() < s->cfg.max_batch) {
				hj_sreq &r = s->reqs[s->queue.front()];
				if (r.gathered != g) break;
				if (fp + r.footprint + 64 * 256 > sl.arena_bytes) break;  // FIFO: wait for memory
				if (!g) {
					const serve_req_plan p = serve_plan_req(s, r);
					if (!ids.empty() && hist_entries + p.hist_entries + p.npairs > (uint64_t)SCAN_CH * 60000) break;
					hist_entries += p.hist_entries + p.npairs;
				}
				fp += r.footprint + 64 * 256;
				ids.push_back(s->queue.front());
				s->queue.pop_front();
			}
			if (ids.empty()) continue;
			for (uint64_t id : ids) s->reqs[id].state = 1;
			const int rc = g ? serve_launch_g(s, sl, ids) : serve_launch(s, sl, ids);
			sl.big = false;
			for (uint64_t id : ids) {
				const hj_sreq &r = s->reqs[id];
				sl.big |= r.R.tuple_num + r.S.tuple_num >= kServeBigTuples;
			}
			if (rc != HJ_OK) {  // these requests fail with the launch's status
				for (uint64_t id : ids) {
					hj_sreq &r = s->reqs[id];
					r.state = 2;
					r.err = rc;
					r.t_done = hj_clock::now();
					s->st.completed++;
				}
				continue;
			}
			for (uint64_t id : ids) s->admitted_now += s->reqs[id].footprint;
			s->st.batches++;
			s->st.max_batch_joins = std::max<uint64_t>(s->st.max_batch_joins, ids.size());
			s->st.peak_admitted_bytes = std::max(s->st.peak_admitted_bytes, s->admitted_now);
			busy++;
			progress = true;
		}
		s->st.max_busy_slots = std::max(s->st.max_busy_slots, busy);
		if (progress) s->cv.notify_all();
		if (s->stop && s->queue.empty() && busy == 0) break;
		// While work is in flight, and for 2 ms after the last, the loop polls
		// with a yield (a timed sleep of tens of microseconds lasts about 70 on
		// Linux, and waking from the condition variable costs as much, both
		// straight onto a request's latency); an idle server blocks.
		const hj_clock::time_point now = hj_clock::now();
		if (progress || busy) last_active = now;
		if (!progress) {
			if (busy == 0 && s->queue.empty() && now - last_active > std::chrono::milliseconds(2))
				s->cv.wait_for(lk, std::chrono::milliseconds(5));
			else {
				lk.unlock();
				std::this_thread::yield();
				lk.lock();
			}
		}
	}
}

// A solo lane runs one large request at a time through the single-join
// engine on its own context (stream + scratch); the device overlaps it with
// the batch slots.
static void solo_loop(hj_server *s, hj_solo_lane *lane) {
	(void)hipSetDevice(s->ctx->device);
	std::unique_lock<std::mutex> lk(s->mu);
	for (;;) {
		s->cv.wait(lk, [&] { return s->stop || !s->solo_queue.empty(); });
		if (s->solo_queue.empty()) {
			if (s->stop) break;
			continue;
		}
		const uint64_t id = s->solo_queue.front();
		s->solo_queue.pop_front();
		hj_sreq &r = s->reqs[id];  // stable: unordered_map nodes do not move
		r.state = 1;
		s->solo_busy++;
		const hj_relation_t R = r.R, S = r.S;
		lk.unlock();
		hj_params_t p{};
		hj_result_t out{};
		const int rc = hj_join(lane->ctx, s->algo.c_str(), &R, &S, &p, &out);
		lk.lock();
		r.err = rc;
		r.res.matched_cnt = out.matched_cnt;
		r.res.check_sum = out.check_sum;
		r.t_done = hj_clock::now();
		r.state = 2;
		s->solo_busy--;
		s->st.completed++;
		s->st.solo_completed++;
		s->cv.notify_all();
	}
}

// Scratch of the single-join engine for n = |R| + |S| tuples of W bytes (two
// passes, the worst case): both partition buffers, the pass-A count matrix
// (at most 2 B per tuple: a 4096-tuple tile has at most 1024 first-pass
// digits of 8 B, or 128 digits per 2048-tuple tile below 8 first-pass bits),
// run tables, tile records and work lists (below 1 B per tuple), and a fixed
// allowance for the small buffers.
static uint64_t solo_scratch(uint64_t n, uint32_t W) { return 2 * n * W + 3 * n + (64u << 20); }
static uint64_t solo_footprint(const hj_sreq &r, uint32_t W) { return solo_scratch(r.R.tuple_num + r.S.tuple_num, W); }

// A solo lane's scratch for the largest request it admits, reserved when the
// server opens: a request never waits for the lane's buffers to grow (a grow
// synchronises the lane and reallocates gigabytes).
static int solo_reserve(hj_ctx_t *lc, uint64_t budget, uint32_t W) {
	if (budget <= (64u << 20)) return HJ_OK;
	const uint64_t n = (budget - (64u << 20)) / (2 * (uint64_t)W + 3);
	int rc;
	if ((rc = hj_reserve(lc, lc->part[0], n * W + 4096)) || (rc = hj_reserve(lc, lc->part[1], n * W + 4096)) ||
	    (rc = hj_reserve(lc, lc->hist, 2 * n + 4096)) || (rc = hj_reserve(lc, lc->meta, n + (32u << 20))) ||
	    (rc = hj_reserve(lc, lc->scan_tmp, (2 * n / 8) / SCAN_CH * 8 + (1u << 20))))
		return rc;
	return HJ_OK;
}

extern "C" int hj_serve_open(hj_ctx_t *ctx, const char *algo, const hj_serve_config_t *cfg, hj_server_t **out) {
	if (!ctx || !algo || !cfg || !out) return hj_fail(ctx, HJ_EINVAL, "null argument");
	*out = nullptr;
	bool radix = false;
	const bool lp = hj_algo_is_lp(algo, &radix);
	if (!radix) return hj_fail(ctx, HJ_EINVAL, "serving runs the radix joins (phj_radix_bc/lp), not '%s'", algo);
	if (cfg->slots < 1 || cfg->slots > 8 || cfg->max_batch < 1 || cfg->solo_lanes > 4 || (cfg->flags & ~HJ_SERVE_SCATTER) ||
	    cfg->hbm_cap_bytes < ((uint64_t)(cfg->slots + cfg->solo_lanes) << 24))
		return hj_fail(ctx, HJ_EINVAL, "serve config: 1..8 slots, 0..4 solo lanes, max_batch >= 1, cap >= 16 MiB per lane");
	hj_server *s = new hj_server();
	s->ctx = ctx;
	s->lp = lp;
	s->algo = algo;
	s->cfg = *cfg;
	s->solo_min = 1ull << (cfg->solo_min_log2 ? cfg->solo_min_log2 : 16);
	s->slots.resize(cfg->slots);
	// with solo lanes, half of the cap backs the batch slots, half the lanes
	const uint64_t batch_cap = cfg->solo_lanes ? cfg->hbm_cap_bytes / 2 : cfg->hbm_cap_bytes;
	if (cfg->solo_lanes) s->solo_budget = (cfg->hbm_cap_bytes - batch_cap) / cfg->solo_lanes;
	const size_t per = (size_t)(batch_cap / cfg->slots) & ~(size_t)4095;
	int rc = HJ_OK;
	{
		std::lock_guard<std::mutex> lk(ctx->mu);
		(void)hipSetDevice(ctx->device);
		for (hj_slot &sl : s->slots) {
			if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
			    hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) {
				(void)hipGetLastError();
				rc = hj_fail(ctx, HJ_EHIP, "slot stream/event");
				break;
			}
			if (hipMalloc(&sl.arena, per) != hipSuccess) {
				(void)hipGetLastError();
				sl.arena = nullptr;
				rc = hj_fail(ctx, HJ_ENOMEM, "slot arena of %zu bytes", per);
				break;
			}
			sl.arena_bytes = per;
			s->st.arena_bytes += per;
		}
	}
	if (rc) {
		for (hj_slot &sl : s->slots) {
			if (sl.arena) (void)hipFree(sl.arena);
			if (sl.done) (void)hipEventDestroy(sl.done);
			if (sl.stream) (void)hipStreamDestroy(sl.stream);
		}
		delete s;
		return rc;
	}
	s->solo.resize(cfg->solo_lanes);
	for (hj_solo_lane &l : s->solo) {
		if (hj_ctx_create(ctx->device, &l.ctx) != HJ_OK) {
			for (hj_solo_lane &m : s->solo)
				if (m.ctx) hj_ctx_destroy(m.ctx);
			for (hj_slot &sl : s->slots) {
				(void)hipFree(sl.arena);
				(void)hipEventDestroy(sl.done);
				(void)hipStreamDestroy(sl.stream);
			}
			delete s;
			return hj_fail(ctx, HJ_EHIP, "solo lane context");
		}
		if (ctx->tuned) {  // lanes use the parent's size-class tuning
			memcpy(l.ctx->tune_bits, ctx->tune_bits, sizeof ctx->tune_bits);
			memcpy(l.ctx->tune_passes, ctx->tune_passes, sizeof ctx->tune_passes);
			memcpy(l.ctx->tune_pass1, ctx->tune_pass1, sizeof ctx->tune_pass1);
			memcpy(l.ctx->tune_bits1, ctx->tune_bits1, sizeof ctx->tune_bits1);
			memcpy(l.ctx->tune_var1, ctx->tune_var1, sizeof ctx->tune_var1);
			l.ctx->tuned = true;
		}
		// sized for 16-byte tuples (the width is known at the first request;
		// wider ones may grow the partition buffers once)
		rc = solo_reserve(l.ctx, s->solo_budget, 16);
		if (rc) break;
	}
	// the slots' pinned staging buffers, sized for a full batch's plan
	for (hj_slot &sl : s->slots)
		if (!rc) rc = serve_grow_pinned(s, sl, 8u << 20);
	if (rc) {
		for (hj_solo_lane &m : s->solo)
			if (m.ctx) hj_ctx_destroy(m.ctx);
		for (hj_slot &sl : s->slots) {
			if (sl.hbuf) (void)hipHostFree(sl.hbuf);
			(void)hipFree(sl.arena);
			(void)hipEventDestroy(sl.done);
			(void)hipStreamDestroy(sl.stream);
		}
		delete s;
		return rc;
	}
	s->th = std::thread(serve_loop, s);
	for (hj_solo_lane &l : s->solo) l.th = std::thread(solo_loop, s, &l);
	*out = s;
	return HJ_OK;
}

extern "C" int hj_serve_register_build(hj_server_t *s, const hj_relation_t *R, uint32_t *build_id) {
	if (!s || !R || !build_id) return HJ_EINVAL;
	hj_ctx_t *ctx = s->ctx;
	if (R->location != HJ_LOC_DEVICE || (R->tuple_num && !R->tuples)) return hj_fail(ctx, HJ_EINVAL, "device build relation required");
	std::lock_guard<std::mutex> lk(s->mu);
	if (s->W == 0) s->W = R->tuple_size;
	if (R->tuple_size != s->W || (s->W != 16 && s->W != 32 && s->W != 64 && s->W != 128 && s->W != 256))
		return hj_fail(ctx, HJ_EINVAL, "tuple size %u", R->tuple_size);
	hj_build_res b;
	b.R = *R;
	if (serve_gathered(s) && R->tuple_num <= serve_gmax(s)) {
		// gathered layout: pass A over R alone; the tiles, digit-major run rows
		// and descriptors stay in one device region for the server's life
		hj_gplan P;
		const uint32_t bits = gbatch_bits(ctx, R->tuple_num);
		const uint32_t si = gplan_seg(P, *R, bits);
		gplan_finish(P, (uint32_t)ctx->num_cu);
		const size_t o_st = al256(P.bytes), total = o_st + al256(sizeof(hj_stats));
		std::vector<unsigned char> img(P.bytes - P.o_up);
		std::lock_guard<std::mutex> lk2(ctx->mu);
		if (hipMalloc(&b.gbuf, total) != hipSuccess) {
			(void)hipGetLastError();
			b.gbuf = nullptr;
			return hj_fail(ctx, HJ_ENOMEM, "resident build of %llu tuples", (unsigned long long)R->tuple_num);
		}
		char *base = (char *)b.gbuf;
		gplan_place(P, base, img.data());
		hj_glaunch L;
		int rc = gplan_prepare(ctx, P, base, L);
		if (!rc && !img.empty() && hipMemcpy(base + P.o_up, img.data(), img.size(), hipMemcpyHostToDevice) != hipSuccess)
			rc = hj_fail(ctx, HJ_EHIP, "build plan upload");
		if (!rc) {
			const bool kt = ctx->kernel_times;
			ctx->kernel_times = false;
			gplan_launch(ctx, ctx->stream, L, (hj_stats *)(base + o_st), nullptr, nullptr);
			ctx->kernel_times = kt;
			const hipError_t e = hipStreamSynchronize(ctx->stream);
			if (e != hipSuccess) rc = hj_fail(ctx, HJ_EHIP, "build partition: %s", hipGetErrorString(e));
		}
		if (rc) {
			(void)hipFree(b.gbuf);
			return rc;
		}
		b.gathered = true;
		b.bits = bits;
		b.g.rt = (const uint16_t *)(base + P.o_runt) + P.segs[si].hist_base;
		b.g.desc = (const uint64_t *)(base + P.o_desc) + P.segs[si].chunk_base;
		b.g.nc = P.segs[si].nchunks;
		b.g.bits = bits;
		b.bytes = total;
	}
	// a gathered server's builds above the gathered limit run in its solo
	// lanes, on R itself
	if (!b.gathered && !(serve_gathered(s) && !s->solo.empty()))
		b.bits = R->tuple_num > 2048 ? hj_radix_bits_onepass(ctx, R->tuple_num) : 0;
	if (!b.gathered && b.bits) {
		std::lock_guard<std::mutex> lk2(ctx->mu);
		const uint32_t wv = s->W / 16, F = 1u << b.bits;
		const uint64_t n = R->tuple_num, nch = (n + kServeChunk - 1) / kServeChunk;
		hj_seg g{};
		g.in = (const uint4 *)R->tuples;
		g.len = n;
		g.bits = b.bits;
		g.nchunks = (uint32_t)nch;
		std::vector<uint32_t> cmap(nch, 0);
		char *tmp = nullptr;
		const size_t t_hist = al256((nch * F + 1) * 8), t_scan = al256(((nch * F) / SCAN_CH + 4) * 8 + 4096);
		const size_t t_seg = al256(sizeof(hj_seg)), t_cmap = al256(nch * 4), t_st = al256(sizeof(hj_stats));
		if (hipMalloc(&b.part, n * s->W + 64) != hipSuccess || hipMalloc((void **)&b.pb, (F + 1) * 8) != hipSuccess ||
		    hipMalloc(&tmp, t_hist + t_scan + t_seg + t_cmap + t_st) != hipSuccess) {
			(void)hipGetLastError();
			if (b.part) (void)hipFree(b.part);
			if (b.pb) (void)hipFree(b.pb);
			return hj_fail(ctx, HJ_ENOMEM, "resident build of %llu tuples", (unsigned long long)n);
		}
		uint64_t *d_hist = (uint64_t *)tmp, *d_scan = (uint64_t *)(tmp + t_hist);
		hj_seg *d_seg = (hj_seg *)(tmp + t_hist + t_scan);
		uint32_t *d_cmap = (uint32_t *)(tmp + t_hist + t_scan + t_seg);
		hj_stats *d_st = (hj_stats *)(tmp + t_hist + t_scan + t_seg + t_cmap);
		int rc = HJ_OK;
		if (hipMemcpy(d_seg, &g, sizeof g, hipMemcpyHostToDevice) != hipSuccess ||
		    hipMemcpy(d_cmap, cmap.data(), nch * 4, hipMemcpyHostToDevice) != hipSuccess)
			rc = hj_fail(ctx, HJ_EHIP, "build plan upload");
		if (!rc) {
			const bool kt = ctx->kernel_times;
			ctx->kernel_times = false;
			hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, ctx->stream, d_st, nch, (unsigned long long *)nullptr, 0u,
			                   (unsigned int *)nullptr);
			rc = launch_pass(ctx, wv, 0, (unsigned)nch, d_seg, d_cmap, kServeChunk, 0, b.bits, d_hist, nch * F, b.part,
			                 d_st, d_scan);
			if (!rc) hipLaunchKernelGGL(k_bounds_b, dim3(1), dim3(256), 0, ctx->stream, d_seg, 1u, (const uint64_t *)d_hist, b.pb);
			ctx->kernel_times = kt;
			hipError_t e = hipStreamSynchronize(ctx->stream);
			if (!rc && e != hipSuccess) rc = hj_fail(ctx, HJ_EHIP, "build partition: %s", hipGetErrorString(e));
		}
		(void)hipFree(tmp);
		if (rc) {
			(void)hipFree(b.part);
			(void)hipFree(b.pb);
			return rc;
		}
		b.bytes = n * s->W + 64 + (F + 1) * 8;
	}
	s->builds.push_back(b);
	s->st.n_builds++;
	s->st.resident_bytes += b.bytes;
	*build_id = (uint32_t)s->builds.size() - 1;
	return HJ_OK;
}

extern "C" int hj_serve_submit(hj_server_t *s, int32_t build_id, const hj_relation_t *R, const hj_relation_t *S,
                               uint64_t *req_id) {
	if (!s || !S || !req_id) return HJ_EINVAL;
	hj_ctx_t *ctx = s->ctx;
	std::unique_lock<std::mutex> lk(s->mu);
	hj_sreq r;
	r.build = build_id;
	if (build_id >= 0) {
		if ((size_t)build_id >= s->builds.size()) return hj_fail(ctx, HJ_EINVAL, "unknown build %d", build_id);
		r.R = s->builds[build_id].R;
	} else {
		if (!R) return hj_fail(ctx, HJ_EINVAL, "null build relation");
		r.R = *R;
	}
	r.S = *S;
	if (s->W == 0) s->W = r.R.tuple_size;
	if (r.R.tuple_size != s->W || r.S.tuple_size != s->W) return hj_fail(ctx, HJ_EINVAL, "tuple sizes must match the server's");
	if (s->W != 16 && s->W != 32 && s->W != 64 && s->W != 128 && s->W != 256)
		return hj_fail(ctx, HJ_EINVAL, "tuple size %u", s->W);
	if (r.R.location != HJ_LOC_DEVICE || r.S.location != HJ_LOC_DEVICE || (r.R.tuple_num && !r.R.tuples) ||
	    (r.S.tuple_num && !r.S.tuples))
		return hj_fail(ctx, HJ_EINVAL, "device relations required");
	// gathered servers batch every gatherable request, but send build sides
	// above 2^20 to an idle solo lane; otherwise large build sides
	// (>= solo_min) go to the solo lanes
	r.gathered = serve_req_gatherable(s, r);
	const bool lane_idle = s->solo_busy + s->solo_queue.size() < s->solo.size();
	bool solo = !s->solo.empty() && !r.gathered && (serve_gathered(s) || r.R.tuple_num >= s->solo_min);
	if (r.gathered && lane_idle && r.R.tuple_num > kServeSoloGathered) {
		solo = true;
		r.gathered = false;
	}
	r.footprint = solo ? solo_footprint(r, s->W) : r.gathered ? serve_gfootprint(s, r) : serve_footprint(serve_plan_req(s, r), s->W);
	s->st.submitted++;
	const uint64_t room = solo ? s->solo_budget : s->slots[0].arena_bytes - kServeBatchOverhead - 64 * 256;
	if (r.footprint > room) {
		s->st.rejected++;
		return hj_fail(ctx, HJ_EREJECT, "request needs %llu bytes of scratch, its lane holds %llu",
		               (unsigned long long)r.footprint, (unsigned long long)room);
	}
	r.t_arr = hj_clock::now();
	const uint64_t id = s->next_id++;
	s->reqs.emplace(id, r);
	(solo ? s->solo_queue : s->queue).push_back(id);
	*req_id = id;
	lk.unlock();
	s->cv.notify_all();
	return HJ_OK;
}

extern "C" int hj_serve_poll(hj_server_t *s, uint64_t req_id, hj_join_result_t *res, double *latency_ms) {
	if (!s) return HJ_EINVAL;
	std::lock_guard<std::mutex> lk(s->mu);
	auto it = s->reqs.find(req_id);
	if (it == s->reqs.end()) return hj_fail(s->ctx, HJ_EINVAL, "unknown request %llu", (unsigned long long)req_id);
	if (it->second.state != 2) return HJ_EBUSY;
	if (it->second.err != HJ_OK) return hj_fail(s->ctx, it->second.err, "request %llu failed", (unsigned long long)req_id);
	if (res) *res = it->second.res;
	if (latency_ms)
		*latency_ms = std::chrono::duration<double, std::milli>(it->second.t_done - it->second.t_arr).count();
	return HJ_OK;
}

extern "C" int hj_serve_drain(hj_server_t *s) {
	if (!s) return HJ_EINVAL;
	std::unique_lock<std::mutex> lk(s->mu);
	s->cv.wait(lk, [&] { return s->st.completed + s->st.rejected >= s->st.submitted; });
	return HJ_OK;
}

extern "C" int hj_serve_stats(hj_server_t *s, hj_serve_stats_t *st) {
	if (!s || !st) return HJ_EINVAL;
	std::lock_guard<std::mutex> lk(s->mu);
	*st = s->st;
	return HJ_OK;
}

extern "C" int hj_serve_close(hj_server_t *s) {
	if (!s) return HJ_EINVAL;
	{
		std::lock_guard<std::mutex> lk(s->mu);
		s->stop = true;
	}
	s->cv.notify_all();
	if (s->th.joinable()) s->th.join();
	for (hj_solo_lane &l : s->solo) {
		if (l.th.joinable()) l.th.join();
		if (l.ctx) hj_ctx_destroy(l.ctx);
	}
	(void)hipSetDevice(s->ctx->device);
	for (hj_slot &sl : s->slots) {
		if (sl.stream) (void)hipStreamSynchronize(sl.stream);
		if (sl.arena) (void)hipFree(sl.arena);
		if (sl.hbuf) (void)hipHostFree(sl.hbuf);
		if (sl.done) (void)hipEventDestroy(sl.done);
		if (sl.stream) (void)hipStreamDestroy(sl.stream);
	}
	for (hj_build_res &b : s->builds) {
		if (b.part) (void)hipFree(b.part);
		if (b.pb) (void)hipFree(b.pb);
		if (b.gbuf) (void)hipFree(b.gbuf);
	}
	delete s;
	return HJ_OK;
}
