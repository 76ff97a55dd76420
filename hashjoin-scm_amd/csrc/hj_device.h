// hj_device.h -- device-side helpers: tuple access, wave and block reductions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#define HJ_WAVE 64

// A tuple is WV 16-byte words; word 0 holds {key, payload}.
template <int WV>
struct hj_tuple {
	uint4 w[WV];
};

__device__ __forceinline__ uint64_t hj_key_of(const uint4 &w0) {
	return (uint64_t)w0.x | ((uint64_t)w0.y << 32);
}
__device__ __forceinline__ uint64_t hj_pay_of(const uint4 &w0) {
	return (uint64_t)w0.z | ((uint64_t)w0.w << 32);
}
// 8-byte narrow tuples {u32 key, u32 row_id}
__device__ __forceinline__ uint64_t hj_key_of(const uint2 &t) { return t.x; }
__device__ __forceinline__ uint64_t hj_pay_of(const uint2 &t) { return t.y; }

// Plain (cached) 16-byte load of word 0 of tuple i.
template <int WV>
__device__ __forceinline__ uint4 hj_load_head(const uint4 *__restrict__ base, uint64_t i) {
	return base[i * WV];
}

typedef unsigned int hj_u32x4 __attribute__((ext_vector_type(4)));

// Device memory seen as the global address space.  Pointers read from
// descriptors in memory are generic to the compiler, which then emits FLAT
// instructions; those count on both the vector-memory and the LDS counters
// and may complete out of order with LDS traffic, so every use must wait for
// all outstanding memory operations.  Global instructions keep the counters
// exact, so prefetched loads stay in flight across unrelated work.
#define HJ_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ const HJ_GLOBAL T *hj_g(const T *p) {
	return (const HJ_GLOBAL T *)p;
}
template <typename T>
__device__ __forceinline__ HJ_GLOBAL T *hj_g(T *p) {
	return (HJ_GLOBAL T *)p;
}

// Streaming load hint for data read exactly once.
__device__ __forceinline__ uint4 hj_load_stream(const uint4 *p) {
	const hj_u32x4 v = __builtin_nontemporal_load((const HJ_GLOBAL hj_u32x4 *)p);
	return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void hj_store_stream(uint4 *p, const uint4 &v) {
	hj_u32x4 t = {v.x, v.y, v.z, v.w};
	__builtin_nontemporal_store(t, (HJ_GLOBAL hj_u32x4 *)p);
}
// 8-byte forms for narrow tuples {u32 key; u32 row_id}.
typedef unsigned int hj_u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint2 hj_load_stream8(const uint2 *p) {
	const hj_u32x2 v = __builtin_nontemporal_load((const HJ_GLOBAL hj_u32x2 *)p);
	return make_uint2(v.x, v.y);
}
__device__ __forceinline__ void hj_store_stream8(uint2 *p, uint32_t a, uint32_t b) {
	hj_u32x2 t = {a, b};
	__builtin_nontemporal_store(t, (HJ_GLOBAL hj_u32x2 *)p);
}
// 4-byte form for the positional words of the first radix pass.
__device__ __forceinline__ void hj_store_stream4(uint32_t *p, uint32_t a) {
	__builtin_nontemporal_store(a, (HJ_GLOBAL uint32_t *)p);
}
// 16 and 8 bytes at a 4-byte aligned address (global loads of several
// dwords need dword alignment only), streaming.
typedef unsigned int hj_u32x4a __attribute__((ext_vector_type(4), aligned(4)));
__device__ __forceinline__ uint4 hj_load16a4(const char *p) {
	const hj_u32x4a v = __builtin_nontemporal_load((const HJ_GLOBAL hj_u32x4a *)p);
	return make_uint4(v.x, v.y, v.z, v.w);
}
typedef unsigned int hj_u32x2a __attribute__((ext_vector_type(2), aligned(4)));
__device__ __forceinline__ uint2 hj_load8a4(const char *p) {
	const hj_u32x2a v = __builtin_nontemporal_load((const HJ_GLOBAL hj_u32x2a *)p);
	return make_uint2(v.x, v.y);
}
// cached forms: lines that neighbouring runs share stay in L2 for the
// second reader (a streaming load marks them for early eviction)
__device__ __forceinline__ uint4 hj_load16a4c(const char *p) {
	const hj_u32x4a v = *(const HJ_GLOBAL hj_u32x4a *)p;
	return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint2 hj_load8a4c(const char *p) {
	const hj_u32x2a v = *(const HJ_GLOBAL hj_u32x2a *)p;
	return make_uint2(v.x, v.y);
}
// Plain (cached) global loads.
__device__ __forceinline__ uint4 hj_load(const uint4 *p) {
	const hj_u32x4 v = *(const HJ_GLOBAL hj_u32x4 *)p;
	return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint2 hj_load8(const uint2 *p) {
	const hj_u32x2 v = *(const HJ_GLOBAL hj_u32x2 *)p;
	return make_uint2(v.x, v.y);
}

// Match sinks of the probe loops.  hj_sum_sink counts matches and adds the
// two payloads (the reference's matched_cnt / check_sum).  hj_pair_sink also
// appends the pair (build payload, probe payload) -- the row-id pairs of late
// materialisation -- to a device array.
//
// Pair slots come in per-wave chunks of `chunk` slots claimed with one
// global atomic on `cursor`, so a join of M pairs issues about M / chunk
// atomics (one per match call on a single address serialises the device).
// A wave's chunk (pos, end) lives in LDS (`ws`, two words per wave) because
// matches arrive in divergent code, where lane registers of inactive lanes
// would go stale.  At the end each wave fills the unused rest of its chunk
// with HJ_PAIR_HOLE, which readers skip.  Slots at or past `cap` are not
// written; the cursor then exceeds cap and the caller re-runs with more.
#define HJ_PAIR_HOLE 0xFFFFFFFFFFFFFFFFULL

struct hj_sum_sink {
	unsigned long long cnt = 0, sum = 0;
	__device__ __forceinline__ void match(uint64_t r, uint64_t s) {
		cnt++;
		sum += r + s;
	}
	__device__ __forceinline__ void finish() {}
};

struct hj_pair_sink {
	unsigned long long cnt = 0, sum = 0;
	uint4 *out = nullptr;
	unsigned long long cap = 0;
	unsigned long long *cursor = nullptr;
	volatile unsigned long long *ws = nullptr;  // LDS: pos, end per wave
	unsigned chunk = 0;                          // >= 64

	__device__ __forceinline__ void put(unsigned long long k, uint64_t r, uint64_t s) const {
		if (k < cap) {
			hj_u32x4 v = {(uint32_t)r, (uint32_t)(r >> 32), (uint32_t)s, (uint32_t)(s >> 32)};
			__builtin_nontemporal_store(v, (HJ_GLOBAL hj_u32x4 *)(out + k));
		}
	}
	__device__ __forceinline__ void match(uint64_t r, uint64_t s) {
		cnt++;
		sum += r + s;
		const unsigned long long act = __ballot(1);  // lanes storing a pair now
		const int lane = threadIdx.x & (HJ_WAVE - 1), lead = __ffsll((long long)act) - 1;
		const unsigned c = __popcll(act), rank = __popcll(act & ((1ull << lane) - 1));
		volatile unsigned long long *w = ws + 2 * (threadIdx.x / HJ_WAVE);
		const unsigned long long pos = w[0], end = w[1], left = end - pos;
		unsigned long long slot, npos, nend;
		if (c > left) {  // wave-uniform: chunk exhausted
			unsigned long long base = 0;
			if (lane == lead) base = atomicAdd(cursor, (unsigned long long)chunk);
			base = __shfl(base, lead, HJ_WAVE);
			slot = rank < left ? pos + rank : base + (rank - left);
			npos = base + (c - left);
			nend = base + chunk;
		} else {
			slot = pos + rank;
			npos = pos + c;
			nend = end;
		}
		if (lane == lead) {
			w[0] = npos;
			w[1] = nend;
		}
		put(slot, r, s);
	}
	// every lane of the wave, converged: mark the rest of the chunk
	__device__ __forceinline__ void finish() {
		volatile unsigned long long *w = ws + 2 * (threadIdx.x / HJ_WAVE);
		const unsigned long long pos = w[0], end = w[1];
		for (unsigned long long k = pos + (threadIdx.x & (HJ_WAVE - 1)); k < end; k += HJ_WAVE)
			put(k, HJ_PAIR_HOLE, HJ_PAIR_HOLE);
	}
};

// The sink of a probe kernel: hj_pair_sink when EMIT, else hj_sum_sink.
// `ws` is the kernel's LDS array of two words per wave; every lane calls
// this before any match (lane 0 of each wave clears its chunk).
template <bool EMIT>
__device__ __forceinline__ typename std::conditional<EMIT, hj_pair_sink, hj_sum_sink>::type
hj_make_sink(uint4 *out, unsigned long long cap, unsigned long long *cursor, unsigned chunk,
             unsigned long long *ws) {
	typename std::conditional<EMIT, hj_pair_sink, hj_sum_sink>::type acc;
	if constexpr (EMIT) {
		acc.out = out;
		acc.cap = cap;
		acc.cursor = cursor;
		acc.chunk = chunk;
		acc.ws = ws;
		if ((threadIdx.x & (HJ_WAVE - 1)) == 0) {
			acc.ws[2 * (threadIdx.x / HJ_WAVE)] = 0;
			acc.ws[2 * (threadIdx.x / HJ_WAVE) + 1] = 0;
		}
	}
	return acc;
}

__device__ __forceinline__ unsigned long long hj_wave_sum(unsigned long long v) {
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, HJ_WAVE);
	return v;
}

// Reduce (a, b) over the block and add the totals to dst[0], dst[1] with one
// 64-bit atomic each (integer sums mod 2^64: any order gives the same bits).
// All threads of the block must call it; `scratch` holds 2 * waves entries.
__device__ __forceinline__ void hj_block_add2(unsigned long long a, unsigned long long b,
                                              unsigned long long *scratch,
                                              unsigned long long *dst0, unsigned long long *dst1) {
	a = hj_wave_sum(a);
	b = hj_wave_sum(b);
	const int lane = threadIdx.x & (HJ_WAVE - 1);
	const int wave = threadIdx.x / HJ_WAVE;
	const int nw = (blockDim.x + HJ_WAVE - 1) / HJ_WAVE;
	__syncthreads();
	if (lane == 0) {
		scratch[2 * wave] = a;
		scratch[2 * wave + 1] = b;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		unsigned long long ta = 0, tb = 0;
		for (int w = 0; w < nw; w++) {
			ta += scratch[2 * w];
			tb += scratch[2 * w + 1];
		}
		if (ta) atomicAdd(dst0, ta);
		if (tb) atomicAdd(dst1, tb);
	}
	__syncthreads();
}

// Exclusive scan of `v` (one value per thread) across the block; returns the
// exclusive prefix and writes the block total to *total.  `scratch` holds
// (waves + 1) entries.  All threads must call it.
template <typename T>
__device__ __forceinline__ T hj_block_exscan(T v, T *scratch, T *total) {
	const int lane = threadIdx.x & (HJ_WAVE - 1);
	const int wave = threadIdx.x / HJ_WAVE;
	const int nw = (blockDim.x + HJ_WAVE - 1) / HJ_WAVE;
	T incl = v;
#pragma unroll
	for (int off = 1; off < HJ_WAVE; off <<= 1) {
		T t = __shfl_up(incl, off, HJ_WAVE);
		if (lane >= off) incl += t;
	}
	if (lane == HJ_WAVE - 1) scratch[wave] = incl;
	__syncthreads();
	if (threadIdx.x == 0) {
		T run = 0;
		for (int w = 0; w < nw; w++) {
			T t = scratch[w];
			scratch[w] = run;
			run += t;
		}
		scratch[nw] = run;
	}
	__syncthreads();
	T res = scratch[wave] + incl - v;
	*total = scratch[nw];
	__syncthreads();
	return res;
}
