// hj_tables.h -- per-partition LDS hash tables of the radix joins.
//
// A table holds up to `cap` build tuples of one partition (a "fill"); its
// kind follows the reference's per-partition joins (src/algo/phj.c):
//   HJ_TK_BC  bc_join  :496-571  heads u32[nb] + next u16[cap], 1-based,
//                                head insertion, all pairs
//   HJ_TK_LP  lp_join  :574-680  nb slots (index+1 of the tuple), linear
//                                probing, first match only
//   HJ_TK_SC  sc_join  :683-781  nb/2 buckets of 2 inline tuples + an
//                                overflow chain per bucket, all pairs
//   HJ_TK_HR  hr_join  :784-877  tuples counting-sorted by bucket (nb
//                                buckets), a probe scans its bucket's range
//   HJ_TK_HRO hro_join :881-1037 hr with nb >> f buckets (f = log2(64 / W))
//                                and the range read two keys per LDS load
// nb = next power of two of the tuples the fill may hold.  The bucket of a
// key is (key >> B) & (buckets - 1) (IDHASH, params.h:9-11).  All kinds but
// LP count every pair, so their results are identical and a partition may be
// joined in several fills.
//
// LDS layout (u64 arrays first, so every array stays aligned):
//   keys u64[cap] | pays u64[cap] | kind-specific part
//     BC : heads u32[nb] | next u16[cap]
//     LP : slots u32[nb]
//     SC : bkey u64[2*nbk] | bpay u64[2*nbk] | bcnt u32[nbk] | ovf u32[nbk] |
//          next u16[cap]            (keys/pays hold the overflow tuples)
//     HR : hist u32[nb]             (after the build: hist[h] = end of bucket h)
//     HRO: hist u32[nbo]
#pragma once

#include <stdint.h>

#include "hj_device.h"

enum hj_tk : int { HJ_TK_BC = 0, HJ_TK_LP = 1, HJ_TK_SC = 2, HJ_TK_HR = 3, HJ_TK_HRO = 4 };

static inline __host__ __device__ uint32_t hj_np2u(uint32_t x) {
	uint32_t p = 1;
	while (p < x) p <<= 1;
	return p;
}

// hro's bucket shrink: log2(64 / W) for W <= 64 (phj.h:22-24)
static inline __host__ __device__ uint32_t hj_hro_shift(uint32_t wv) { return wv == 1 ? 2 : (wv == 2 ? 1 : 0); }

// buckets of a table kind for a fill of up to n tuples
static inline __host__ __device__ uint32_t hj_tab_buckets(int tk, uint32_t n, uint32_t wv) {
	const uint32_t p = hj_np2u(n ? n : 1);
	if (tk == HJ_TK_SC) return p >= 2 ? p / 2 : 1;
	if (tk == HJ_TK_HRO) {
		const uint32_t f = hj_hro_shift(wv), b = p >> f;
		return b < (1u << f) ? (1u << f) : b;
	}
	return p;
}

// LDS bytes of a table kind for `cap` tuples
static inline __host__ __device__ size_t hj_tab_lds(int tk, uint32_t cap, uint32_t wv) {
	const size_t base = (size_t)cap * 16, nb = hj_tab_buckets(tk, cap, wv);
	switch (tk) {
	case HJ_TK_LP:
	case HJ_TK_HR:
	case HJ_TK_HRO: return base + nb * 4;
	case HJ_TK_SC: return base + nb * 32 + nb * 8 + (size_t)cap * 2 + 8;
	default: return base + nb * 4 + (size_t)cap * 2 + 8;
	}
}

// Device view of one table; reset() picks the fill's bucket count.
template <int TK, int WV>
struct hj_tab {
	uint64_t *keys, *pays;
	uint32_t *heads;        // BC heads / LP slots / SC bucket counts / HR, HRO hist
	uint16_t *next;         // BC next / SC overflow next
	uint64_t *bkey, *bpay;  // SC inline tuples
	uint32_t *ovf;          // SC overflow chain heads (1-based)
	uint32_t *pool;         // SC overflow tuples in use
	uint32_t nb, hmask;

	__device__ void map(uint64_t *lds, uint32_t cap, uint32_t *pool_counter) {
		keys = lds;
		pays = keys + cap;
		pool = pool_counter;
		const uint32_t nbc = hj_tab_buckets(TK, cap, WV);
		if (TK == HJ_TK_SC) {
			bkey = pays + cap;
			bpay = bkey + 2 * nbc;
			heads = (uint32_t *)(bpay + 2 * nbc);
			ovf = heads + nbc;
			next = (uint16_t *)(ovf + nbc);
		} else {
			heads = (uint32_t *)(pays + cap);
			next = (uint16_t *)(heads + nbc);
		}
	}
	// bucket count for a fill of up to n tuples and cleared buckets (all
	// threads; the caller syncs before inserting)
	__device__ void reset(uint32_t n, uint32_t nt) {
		nb = hj_tab_buckets(TK, n, WV);
		hmask = nb - 1;
		for (uint32_t i = threadIdx.x; i < nb; i += nt) {
			heads[i] = 0;
			if (TK == HJ_TK_SC) ovf[i] = 0;
		}
		if (TK == HJ_TK_SC && threadIdx.x == 0) *pool = 0;
	}
	__device__ uint32_t bucket(uint64_t key, uint32_t B) const { return (uint32_t)(key >> B) & hmask; }

	// one-stage builds (BC, SC) and the gather of LP: tuple i of the fill
	__device__ void put(uint32_t i, uint64_t key, uint64_t pay, uint32_t B) {
		if (TK == HJ_TK_SC) {
			const uint32_t h = bucket(key, B), slot = atomicAdd(&heads[h], 1u);
			if (slot < 2) {
				bkey[2 * h + slot] = key;
				bpay[2 * h + slot] = pay;
			} else {
				const uint32_t o = atomicAdd(pool, 1u);
				keys[o] = key;
				pays[o] = pay;
				next[o] = (uint16_t)atomicExch(&ovf[h], o + 1);
			}
			return;
		}
		keys[i] = key;
		pays[i] = pay;
		if (TK == HJ_TK_BC) next[i] = (uint16_t)atomicExch(&heads[bucket(key, B)], i + 1);
	}
	// HR / HRO build, first gather: count per bucket
	__device__ void count(uint64_t key, uint32_t B) { atomicAdd(&heads[bucket(key, B)], 1u); }
	// between the gathers: counts -> bucket starts (all threads)
	__device__ void starts(uint32_t *red, uint32_t nt) {
		const uint32_t per = (nb + nt - 1) / nt, p0 = threadIdx.x * per;
		uint32_t sum = 0;
		for (uint32_t q = 0; q < per && p0 + q < nb; q++) sum += heads[p0 + q];
		uint32_t tot;
		uint32_t run = hj_block_exscan<uint32_t>(sum, red, &tot);
		for (uint32_t q = 0; q < per && p0 + q < nb; q++) {
			const uint32_t c = heads[p0 + q];
			heads[p0 + q] = run;
			run += c;
		}
	}
	// HR / HRO build, second gather: place in bucket order; heads[h] ends as
	// the end of bucket h
	__device__ void place(uint64_t key, uint64_t pay, uint32_t B) {
		const uint32_t pos = atomicAdd(&heads[bucket(key, B)], 1u);
		keys[pos] = key;
		pays[pos] = pay;
	}
	// LP build after the gather of m tuples: claim slots in gather order of
	// arrival; a second build tuple with the same key sets dup
	__device__ void lp_insert(uint32_t m, uint32_t B, uint32_t nt, bool &dup) {
		for (uint32_t i = threadIdx.x; i < m; i += nt) {
			const uint64_t key = keys[i];
			uint32_t p = bucket(key, B);
			for (uint32_t step = 0; step < nb; step++) {
				const uint32_t old = atomicCAS(&heads[p], 0u, i + 1);
				if (old == 0) break;
				if (keys[old - 1] == key) {
					dup = true;
					break;
				}
				p = (p + 1) & hmask;
			}
		}
	}

	// LP: is `key` in the table?  (cross-fill duplicate check)
	__device__ bool lp_has(uint64_t key, uint32_t B) const {
		uint32_t p = bucket(key, B);
		for (uint32_t step = 0; step < nb; step++) {
			const uint32_t o = heads[p];
			if (!o) return false;
			if (keys[o - 1] == key) return true;
			p = (p + 1) & hmask;
		}
		return false;
	}

	// Probe RU keys at once (independent LDS reads overlap); adds matches.
	template <int RU, class Sink>
	__device__ void probe(const uint64_t (&kk)[RU], const uint64_t (&pp)[RU], const bool (&live)[RU], uint32_t B,
	                      Sink &sink) const {
		uint32_t cur[RU], end[RU];
		if (TK == HJ_TK_BC || TK == HJ_TK_SC) {
			if (TK == HJ_TK_SC) {
#pragma unroll
				for (int u = 0; u < RU; u++) {
					if (!live[u]) continue;
					const uint32_t h = bucket(kk[u], B), n = min(heads[h], 2u);
					for (uint32_t s = 0; s < n; s++)
						if (bkey[2 * h + s] == kk[u]) sink.match(bpay[2 * h + s], pp[u]);
				}
			}
			// chain cursors (1-based entries of keys/pays, 0 = end)
#pragma unroll
			for (int u = 0; u < RU; u++)
				cur[u] = live[u] ? (TK == HJ_TK_BC ? heads[bucket(kk[u], B)] : ovf[bucket(kk[u], B)]) : 0u;
			bool any = true;
			while (any) {
				any = false;
#pragma unroll
				for (int u = 0; u < RU; u++) {
					if (!cur[u]) continue;
					const uint32_t i = cur[u] - 1;
					if (keys[i] == kk[u]) sink.match(pays[i], pp[u]);
					cur[u] = next[i];
					any |= cur[u] != 0;
				}
			}
		} else if (TK == HJ_TK_LP) {
			bool act[RU];
#pragma unroll
			for (int u = 0; u < RU; u++) {
				act[u] = live[u];
				cur[u] = bucket(kk[u], B);
			}
			for (uint32_t step = 0; step < nb; step++) {
				bool any = false;
#pragma unroll
				for (int u = 0; u < RU; u++) {
					if (!act[u]) continue;
					const uint32_t o = heads[cur[u]];
					if (!o) {
						act[u] = false;
					} else if (keys[o - 1] == kk[u]) {
						sink.match(pays[o - 1], pp[u]);
						act[u] = false;
					} else {
						cur[u] = (cur[u] + 1) & hmask;
					}
					any |= act[u];
				}
				if (!any) break;
			}
		} else {  // HR, HRO: scan the bucket's range
			uint32_t beg[RU];
#pragma unroll
			for (int u = 0; u < RU; u++) {
				const uint32_t h = bucket(kk[u], B);
				beg[u] = live[u] ? (h ? heads[h - 1] : 0u) : 0u;
				end[u] = live[u] ? heads[h] : 0u;
				cur[u] = TK == HJ_TK_HRO ? beg[u] & ~1u : beg[u];  // hro: two keys per 16-byte read
			}
			bool any = true;
			while (any) {
				any = false;
#pragma unroll
				for (int u = 0; u < RU; u++) {
					if (cur[u] >= end[u]) continue;
					if (TK == HJ_TK_HR) {
						if (keys[cur[u]] == kk[u]) sink.match(pays[cur[u]], pp[u]);
						cur[u]++;
					} else {
						// keys[lo] and keys[lo + 1] in one read; a key before
						// the range (odd start) or past its end is skipped
						const uint32_t lo = cur[u];
						const uint4 two = *(const uint4 *)&keys[lo];
						if (lo >= beg[u] && hj_key_of(two) == kk[u]) sink.match(pays[lo], pp[u]);
						if (lo + 1 < end[u] && hj_pay_of(two) == kk[u]) sink.match(pays[lo + 1], pp[u]);
						cur[u] += 2;
					}
					any |= cur[u] < end[u];
				}
			}
		}
	}
};
