// hj_tpch.hip -- TPC-H Q14 and Q19 on the join engine: the reference's
// end-to-end consumer (src/tpch/tpch_query.c, src/tpch/algo/tpch_nphj.c,
// src/tpch/algo/tpch_rdx.c, src/tpch/inc/tpch_utils.c).
//
// The reference pushes the selections down (part_filter, lineitem_filter,
// tpch_utils.c:44-112), joins part and lineitem on partkey, and sums
// l_extendedprice * (1 - l_discount) over the matching pairs: Q14 splits the
// sum by p_type LIKE 'PROMO%' (tpch_nphj.c:302-309), Q19 keeps the pairs that
// satisfy one of three brand / container / size / quantity groups
// (match_filter, tpch_utils.c:115-160; tpch_nphj.c:310-319).
//
// Here the sums come out of the join itself.  Every selected lineitem becomes
// a 16-byte {l_partkey, revenue} tuple and every selected part a
// {p_partkey, 0} tuple, so the join's check_sum -- the sum of build plus probe
// payload over the matching pairs -- is the revenue sum, exactly, in
// integers: prices are cents and discounts hundredths, so a revenue is
// cents * (100 - discount) in units of 1e-4.  Q14 runs two joins (all parts,
// PROMO parts) against the September 1995 lineitems.  Q19 runs one join per
// group, each against the lineitems whose quantity lies in that group's
// range: the groups' brands differ, so a part is in at most one group and
// every pair that passes match_filter is counted once.  The selections are one
// pass over the columns: each workgroup writes the offsets of its selected rows
// into its own region of a per-output list (counts kept in registers, no global
// atomic), a scan of the per-workgroup counts places the regions, and a gather
// builds the join inputs.
//
// Columns are structure-of-arrays in HBM; strings are the reference's
// fixed-width, NUL-padded char arrays (schema.h:37-67).  A seeded generator
// fills them with the value domains of the TPC-H specification (not dbgen's
// random streams); readers take dbgen's .tbl files and the reference's packed
// part.bin / lineitem.bin (tpch_tbl2bin.c:23-150).

#include "hj_device.h"
#include "hj_internal.h"
#include "hjgen.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#define TP_THREADS 256

namespace {

// ---------------------------------------------------------------- strings
// strncmp(s, lit, width) == 0 for a literal shorter than width
__device__ __forceinline__ bool str_eq(const char *s, const char *lit, int width) {
	int i = 0;
	for (; lit[i] && i < width; i++)
		if (s[i] != lit[i]) return false;
	return i == width || s[i] == 0;
}
// strncmp(s, lit, strlen(lit)) == 0
__device__ __forceinline__ bool str_prefix(const char *s, const char *lit) {
	for (int i = 0; lit[i]; i++)
		if (s[i] != lit[i]) return false;
	return true;
}
// fast_atoi (src/inc/utils.h:185-191) on the digits of a field of `width`
// bytes starting at s
__device__ __forceinline__ int field_atoi(const char *s, int width) {
	int v = 0;
	for (int i = 0; i < width && s[i] >= '0' && s[i] <= '9'; i++) v = v * 10 + (s[i] - '0');
	return v;
}

// ---------------------------------------------------------------- selections
// A selection pass works on tiles of rows.  The string columns a predicate
// tests for every row are staged in LDS for the whole tile with coalesced
// 16-byte loads (a byte-by-byte compare straight from HBM would serialise one
// load latency per character); the loads of tile t + 1 are issued into
// registers before tile t is evaluated, so their latency hides behind the
// work.  Every thread then tests its rows from LDS; the other columns are read
// straight from HBM, and only for rows still in the running.

// One thread's share of a tile's bytes [t0 * W, (t0 + rows) * W) of a column
// of W-byte strings, in registers.  Byte j of the range goes to LDS byte
// head + j, head = the range's address mod 16; thread x holds chunks
// x, x + NT, ... of 16 bytes.  Every load is a whole aligned 16-byte chunk;
// the first and last chunk may hold bytes outside the tile, which no reader
// looks at, and an aligned chunk holding a byte of the column lies in that
// byte's page, so no load leaves mapped memory.
template <int W, int ROWS>
struct stager {
	static constexpr uint32_t NCH = ((uint32_t)ROWS * W + 15) / 16 + 1;  // most chunks a tile spans
	static constexpr int CH = (int)((NCH + TP_THREADS - 1) / TP_THREADS);
	static constexpr size_t LDS = (size_t)NCH * 16;
	uint4 v[CH];
	uint32_t head, nch;
	__device__ __forceinline__ void load(const char *col, uint64_t t0, uint32_t rows) {
		const char *g = col + t0 * W;
		head = (uint32_t)((uintptr_t)g & 15);
		const uint4 *g16 = (const uint4 *)(g - head);
		const uint32_t end = head + rows * W;
		nch = (end + 15) / 16;
		// Whole aligned 16-byte chunks only, every thread the same number of
		// loads (past nch: the last chunk again), none under a lane
		// condition, so no load waits for the next one's issue.  The first
		// and last chunk may hold bytes outside the tile; readers never look
		// at them, and an aligned 16-byte chunk holding at least one byte of
		// the column lies in that byte's page, so the load cannot fault.
		if (nch == 0) return;  // an empty tile at an aligned address
#pragma unroll
		for (int k = 0; k < CH; k++) {
			const uint32_t c = min(threadIdx.x + (uint32_t)k * TP_THREADS, nch - 1);
			v[k] = hj_load_stream(g16 + c);
		}
	}
	// write the chunks to `lds` (16-byte aligned, LDS bytes); returns head
	__device__ __forceinline__ uint32_t store(char *lds) const {
#pragma unroll
		for (int k = 0; k < CH; k++) {
			const uint32_t c = threadIdx.x + (uint32_t)k * TP_THREADS;
			if (c < nch) ((uint4 *)lds)[c] = v[k];
		}
		return head;
	}
};

// The 8 bytes at s as one big-endian integer: comparing two of these is
// strncmp(.., 8) whenever the literal has no NUL in its first 8 bytes.
__device__ __forceinline__ uint64_t be8(const char *s) {
	uint64_t v = 0;
#pragma unroll
	for (int i = 0; i < 8; i++) v = v << 8 | (unsigned char)s[i];
	return v;
}
__host__ __device__ constexpr uint64_t be8_lit(const char *s) {
	uint64_t v = 0;
	for (int i = 0; i < 8; i++) v = v << 8 | (unsigned char)s[i];
	return v;
}

// Q14 lineitem_filter (tpch_utils.c:89-94): strncmp(l_shipdate, "1995-09-01",
// 8) >= 0 and strncmp(l_shipdate, "1995-10-01", 8) < 0
__device__ __forceinline__ bool q14_lineitem(const char *d) {
	const uint64_t v = be8(d);
	return v >= be8_lit("1995-09-01") && v < be8_lit("1995-10-01");
}

// Q19 part_filter (tpch_utils.c:44-82) as a group index: 0, 1, 2 or -1.
// brand = fast_atoi(p_brand + 6)
__device__ __forceinline__ int q19_part_group(const char *brand, int size, const char *ct) {
	const int b = field_atoi(brand + 6, 4);
	if (b == 12 && size >= 1 && size <= 5 &&
	    (str_eq(ct, "SM CASE", 10) || str_eq(ct, "SM BOX", 10) || str_eq(ct, "SM PACK", 10) || str_eq(ct, "SM PKG", 10)))
		return 0;
	if (b == 23 && size >= 1 && size <= 10 &&
	    (str_eq(ct, "MED BAG", 10) || str_eq(ct, "MED BOX", 10) || str_eq(ct, "MED PKG", 10) ||
	     str_eq(ct, "MED PACK", 10)))
		return 1;
	if (b == 34 && size >= 1 && size <= 15 &&
	    (str_eq(ct, "LG CASE", 10) || str_eq(ct, "LG BOX", 10) || str_eq(ct, "LG PACK", 10) || str_eq(ct, "LG PKG", 10)))
		return 2;
	return -1;
}

// Q19 lineitem_filter (tpch_utils.c:95-108) as the set of groups whose
// quantity range holds l_quantity (the ranges 1-11, 10-20, 20-30 overlap).
// The shipping mode literals are the reference's: "AIR" or "AIR REG".
__device__ __forceinline__ unsigned q19_quantity_groups(int q) {
	return (q >= 1 && q <= 11 ? 1u : 0u) | (q >= 10 && q <= 20 ? 2u : 0u) | (q >= 20 && q <= 30 ? 4u : 0u);
}

// revenue = cents * (100 - discount hundredths), in units of 1e-4
__device__ __forceinline__ uint64_t revenue(int64_t cents, int32_t disc) { return (uint64_t)(cents * (100 - disc)); }

// Row selectors.  ITEMS rows per thread and tile.  pre = the registers that
// carry tile t + 1 while tile t is tested (staged string columns plus the
// per-row numbers a test needs); cur = what a test reads of the current tile
// besides LDS; LDS = the staged columns' bytes.  load() fetches a tile into
// pre, store() moves it to LDS and cur, test(r, u, lds, cur) returns the
// outputs (bit o = output o) the tile's row r (item u of its thread) goes to,
// and emit(i) builds row i's {key, payload} tuple.
struct sel_q14_parts {  // all parts -> output 0; PROMO parts -> output 1 too
	static constexpr int ITEMS = 4;
	using st_type = stager<25, ITEMS * TP_THREADS>;
	struct pre {
		st_type type;
	};
	struct cur {
		uint32_t h;
	};
	static constexpr size_t LDS = st_type::LDS;
	const uint64_t *partkey;
	const char *type;
	__device__ void load(pre &p, uint64_t t0, uint32_t rows) const { p.type.load(type, t0, rows); }
	__device__ void store(const pre &p, char *lds, cur &c) const { c.h = p.type.store(lds); }
	__device__ unsigned test(uint32_t r, int, const char *lds, const cur &c) const {
		return 1u | (str_prefix(lds + c.h + r * 25, "PROMO") ? 2u : 0u);
	}
	__device__ uint4 emit(uint64_t i) const { return make_uint4((uint32_t)partkey[i], (uint32_t)(partkey[i] >> 32), 0, 0); }
};
struct sel_q14_lineitems {
	static constexpr int ITEMS = 8;
	using st_date = stager<10, ITEMS * TP_THREADS>;
	struct pre {
		st_date date;
	};
	struct cur {
		uint32_t h;
	};
	static constexpr size_t LDS = st_date::LDS;
	const uint64_t *partkey;
	const int64_t *price;
	const int32_t *disc;
	const char *shipdate;
	__device__ void load(pre &p, uint64_t t0, uint32_t rows) const { p.date.load(shipdate, t0, rows); }
	__device__ void store(const pre &p, char *lds, cur &c) const { c.h = p.date.store(lds); }
	__device__ unsigned test(uint32_t r, int, const char *lds, const cur &c) const {
		return q14_lineitem(lds + c.h + r * 10) ? 1u : 0u;
	}
	__device__ uint4 emit(uint64_t i) const {
		const uint64_t k = partkey[i], p = revenue(price[i], disc[i]);
		return make_uint4((uint32_t)k, (uint32_t)(k >> 32), (uint32_t)p, (uint32_t)(p >> 32));
	}
};
struct sel_q19_parts {
	static constexpr int ITEMS = 4;
	using st10 = stager<10, ITEMS * TP_THREADS>;
	struct pre {
		st10 brand, container;
		int32_t size[ITEMS];
	};
	struct cur {
		uint32_t hb, hc;
		int32_t size[ITEMS];
	};
	static constexpr size_t LDS = 2 * st10::LDS;
	const uint64_t *partkey;
	const char *brand;
	const int32_t *size;
	const char *container;
	__device__ void load(pre &p, uint64_t t0, uint32_t rows) const {
		p.brand.load(brand, t0, rows);
		p.container.load(container, t0, rows);
#pragma unroll
		for (int u = 0; u < ITEMS; u++) {
			const uint32_t r = (uint32_t)u * TP_THREADS + threadIdx.x;
			p.size[u] = size[t0 + min(r, rows - 1)];  // clamped, unconditional (rows >= 1); masked at the test
		}
	}
	__device__ void store(const pre &p, char *lds, cur &c) const {
		c.hb = p.brand.store(lds);
		c.hc = p.container.store(lds + st10::LDS);
#pragma unroll
		for (int u = 0; u < ITEMS; u++) c.size[u] = p.size[u];
	}
	__device__ unsigned test(uint32_t r, int u, const char *lds, const cur &c) const {
		const int g = q19_part_group(lds + c.hb + r * 10, c.size[u], lds + st10::LDS + c.hc + r * 10);
		return g < 0 ? 0u : 1u << g;
	}
	__device__ uint4 emit(uint64_t i) const { return make_uint4((uint32_t)partkey[i], (uint32_t)(partkey[i] >> 32), 0, 0); }
};
struct sel_q19_lineitems {
	static constexpr int ITEMS = 4;
	using st_mode = stager<10, ITEMS * TP_THREADS>;
	using st_instr = stager<25, ITEMS * TP_THREADS>;
	struct pre {
		st_mode mode;
		st_instr instr;
		int32_t q[ITEMS];
	};
	struct cur {
		uint32_t hm, hi;
		int32_t q[ITEMS];
	};
	static constexpr size_t LDS = st_mode::LDS + st_instr::LDS;
	const uint64_t *partkey;
	const int32_t *quantity;
	const int64_t *price;
	const int32_t *disc;
	const char *instr;
	const char *mode;
	__device__ void load(pre &p, uint64_t t0, uint32_t rows) const {
		p.mode.load(mode, t0, rows);
		p.instr.load(instr, t0, rows);
#pragma unroll
		for (int u = 0; u < ITEMS; u++) {
			const uint32_t r = (uint32_t)u * TP_THREADS + threadIdx.x;
			p.q[u] = quantity[t0 + min(r, rows - 1)];  // clamped, unconditional (rows >= 1); masked at the test
		}
	}
	__device__ void store(const pre &p, char *lds, cur &c) const {
		c.hm = p.mode.store(lds);
		c.hi = p.instr.store(lds + st_mode::LDS);
#pragma unroll
		for (int u = 0; u < ITEMS; u++) c.q[u] = p.q[u];
	}
	__device__ unsigned test(uint32_t r, int u, const char *lds, const cur &c) const {
		const unsigned m = q19_quantity_groups(c.q[u]);
		const char *md = lds + c.hm + r * 10;
		if (!m || !(str_eq(md, "AIR", 10) || str_eq(md, "AIR REG", 10)) ||
		    !str_eq(lds + st_mode::LDS + c.hi + r * 25, "DELIVER IN PERSON", 25))
			return 0;
		return m;
	}
	__device__ uint4 emit(uint64_t i) const {
		const uint64_t k = partkey[i], p = revenue(price[i], disc[i]);
		return make_uint4((uint32_t)k, (uint32_t)(k >> 32), (uint32_t)p, (uint32_t)(p >> 32));
	}
};

// Select rows for up to three outputs without a global atomic: workgroup b
// takes the contiguous rows [b * per_block, (b + 1) * per_block) and appends
// the offsets (row - b * per_block) of its selected rows, in row order, to its
// own region of each staging list (s_o + b * per_block, room for all of its
// rows), keeping the running counts in registers.  counts[b * 4 + o] receives
// its count of output o (o = 3: rows selected at all).  The tuples themselves
// are built by the gather, so no load here waits on a test.
template <class Sel>
__global__ __launch_bounds__(TP_THREADS) void k_select(uint64_t n, uint64_t per_block, Sel sel, uint32_t *s0,
                                                       uint32_t *s1, uint32_t *s2, unsigned long long *counts) {
	constexpr int ITEMS = Sel::ITEMS;
	extern __shared__ uint4 tp_lds[];
	__shared__ unsigned long long red[TP_THREADS / HJ_WAVE + 1];
	char *lds = (char *)tp_lds;
	const uint64_t tile = (uint64_t)TP_THREADS * ITEMS;
	const uint64_t r0 = (uint64_t)blockIdx.x * per_block, r1 = min(n, r0 + per_block);
	uint64_t run[4] = {0, 0, 0, 0};  // the workgroup's selections so far (same in every thread)
	typename Sel::pre pf;
	typename Sel::cur cs;
	if (r0 < r1) sel.load(pf, r0, (uint32_t)min(tile, r1 - r0));
	for (uint64_t t0 = r0; t0 < r1; t0 += tile) {
		const uint32_t rows = (uint32_t)min(tile, r1 - t0);
		sel.store(pf, lds, cs);
		__syncthreads();
		if (t0 + tile < r1) sel.load(pf, t0 + tile, (uint32_t)min(tile, r1 - t0 - tile));
		unsigned mask[ITEMS];
		// the four per-thread counts in 16-bit fields of one word (a tile
		// has fewer than 2^16 rows), so one scan serves every output
		unsigned long long cnt = 0;
#pragma unroll
		for (int u = 0; u < ITEMS; u++) {
			const uint32_t r = (uint32_t)u * TP_THREADS + threadIdx.x;
			mask[u] = r < rows ? sel.test(r, u, lds, cs) : 0u;
			cnt += (unsigned long long)(mask[u] & 1u) | (unsigned long long)(mask[u] >> 1 & 1u) << 16 |
			       (unsigned long long)(mask[u] >> 2 & 1u) << 32 | (unsigned long long)(mask[u] != 0) << 48;
		}
		unsigned long long tot;
		const unsigned long long x = hj_block_exscan<unsigned long long>(cnt, red, &tot);
		uint64_t at[3];
#pragma unroll
		for (int o = 0; o < 4; o++) {
			if (o < 3) at[o] = r0 + run[o] + (x >> (16 * o) & 0xffffu);
			run[o] += tot >> (16 * o) & 0xffffu;
		}
		const uint32_t base = (uint32_t)(t0 - r0) + threadIdx.x;
#pragma unroll
		for (int u = 0; u < ITEMS; u++) {
			const uint32_t off = base + (uint32_t)u * TP_THREADS;
			if (mask[u] & 1u) s0[at[0]++] = off;
			if (mask[u] & 2u) s1[at[1]++] = off;
			if (mask[u] & 4u) s2[at[2]++] = off;
		}
		__syncthreads();  // the next tile's store to LDS overwrites this one
	}
	if (threadIdx.x == 0)
#pragma unroll
		for (int o = 0; o < 4; o++) counts[(uint64_t)blockIdx.x * 4 + o] = run[o];
}

// One workgroup: counts[b * 4 + o] -> exclusive prefix over b per output o;
// totals[o] = the sums.
__global__ __launch_bounds__(256) void k_select_scan(unsigned long long *counts, uint32_t nblk,
                                                     unsigned long long *totals) {
	__shared__ unsigned long long red[256 / HJ_WAVE + 1];
	for (int o = 0; o < 4; o++) {
		unsigned long long run = 0;
		for (uint32_t b0 = 0; b0 < nblk; b0 += 256) {
			const uint32_t b = b0 + threadIdx.x;
			const unsigned long long v = b < nblk ? counts[(uint64_t)b * 4 + o] : 0;
			unsigned long long tot;
			const unsigned long long x = hj_block_exscan<unsigned long long>(v, red, &tot);
			if (b < nblk) counts[(uint64_t)b * 4 + o] = run + x;
			run += tot;
		}
		if (threadIdx.x == 0) totals[o] = run;
	}
}

// Workgroup (b, o): build the tuples of workgroup b's selections for output o
// and write them to their place in the contiguous output.  Four rows per
// thread and step, their loads issued together.
template <class Sel>
__global__ __launch_bounds__(256) void k_select_gather(Sel sel, const uint32_t *s0, const uint32_t *s1,
                                                       const uint32_t *s2, uint4 *d0, uint4 *d1, uint4 *d2,
                                                       const unsigned long long *prefix,
                                                       const unsigned long long *totals, uint32_t nblk,
                                                       uint64_t per_block) {
	const uint32_t b = blockIdx.x, o = blockIdx.y;
	const uint64_t row0 = (uint64_t)b * per_block;
	const uint32_t *src = (o == 0 ? s0 : o == 1 ? s1 : s2) + row0;
	uint4 *dst = o == 0 ? d0 : o == 1 ? d1 : d2;
	const unsigned long long beg = prefix[(uint64_t)b * 4 + o];
	const unsigned long long cnt = (b + 1 < nblk ? prefix[(uint64_t)(b + 1) * 4 + o] : totals[o]) - beg;
	for (unsigned long long i0 = 0; i0 < cnt; i0 += 4 * 256) {
		// loads at clamped indices (cnt > 0 here), none under a lane
		// condition, so the four rows' loads are in flight together
		uint32_t off[4];
#pragma unroll
		for (int k = 0; k < 4; k++) off[k] = src[min(i0 + (unsigned long long)k * 256 + threadIdx.x, cnt - 1)];
		uint4 t[4];
#pragma unroll
		for (int k = 0; k < 4; k++) t[k] = sel.emit(row0 + off[k]);
#pragma unroll
		for (int k = 0; k < 4; k++) {
			const unsigned long long i = i0 + (unsigned long long)k * 256 + threadIdx.x;
			if (i < cnt) hj_store_stream(dst + beg + i, t[k]);
		}
	}
}

// Rows per workgroup of a selection over n rows: whole tiles, at most
// num_cu * 8 workgroups.
template <class Sel>
uint64_t select_per_block(hj_ctx_t *ctx, uint64_t n, uint32_t *nblk) {
	const uint64_t tile = (uint64_t)TP_THREADS * Sel::ITEMS;
	const uint64_t tiles = std::max<uint64_t>(1, (n + tile - 1) / tile);
	const uint64_t g = std::min<uint64_t>(tiles, (uint64_t)ctx->num_cu * 8);
	const uint64_t per = (tiles + g - 1) / g * tile;
	*nblk = (uint32_t)((std::max<uint64_t>(n, 1) + per - 1) / per);
	return per;
}

template <class Sel>
void launch_select(hj_ctx_t *ctx, uint64_t n, uint64_t per_block, uint32_t nblk, const Sel &sel, uint32_t *const *st,
                   unsigned long long *counts) {
	static_assert(Sel::LDS <= 64 * 1024, "staged columns fit the default LDS allocation");
	hipLaunchKernelGGL(k_select<Sel>, dim3(nblk), dim3(TP_THREADS), Sel::LDS, ctx->stream, n, per_block, sel, st[0],
	                   st[1], st[2], counts);
}

template <class Sel>
void launch_gather(hj_ctx_t *ctx, uint32_t outs, uint64_t per_block, uint32_t nblk, const Sel &sel,
                   uint32_t *const *st, uint4 *const *dst, const unsigned long long *prefix,
                   const unsigned long long *totals) {
	hipLaunchKernelGGL(k_select_gather<Sel>, dim3(nblk, outs), dim3(256), 0, ctx->stream, sel, st[0], st[1], st[2],
	                   dst[0], dst[1], dst[2], prefix, totals, nblk, per_block);
}

// ---------------------------------------------------------------- generator
// Value domains of the TPC-H specification (clauses 4.2.2.13 and 4.2.3) drawn
// with this library's counter-based generator: row r's values are a pure
// function of (seed, r), so the device fills any scale factor in one pass.
__constant__ char k_type1[6][9] = {"STANDARD", "SMALL", "MEDIUM", "LARGE", "ECONOMY", "PROMO"};
__constant__ char k_type2[5][10] = {"ANODIZED", "BURNISHED", "PLATED", "POLISHED", "BRUSHED"};
__constant__ char k_type3[5][7] = {"TIN", "NICKEL", "BRASS", "STEEL", "COPPER"};
__constant__ char k_cont1[5][6] = {"SM", "LG", "MED", "JUMBO", "WRAP"};
__constant__ char k_cont2[8][5] = {"CASE", "BOX", "BAG", "JAR", "PKG", "PACK", "CAN", "DRUM"};
__constant__ char k_instr[4][18] = {"DELIVER IN PERSON", "COLLECT COD", "NONE", "TAKE BACK RETURN"};
__constant__ char k_mode[7][8] = {"REG AIR", "AIR", "RAIL", "SHIP", "TRUCK", "MAIL", "FOB"};

#define TPCH_STREAM_PART 21u
#define TPCH_STREAM_LINE 22u

__device__ __forceinline__ uint32_t pick(uint64_t seed, uint32_t stream, uint64_t row, uint32_t field, uint32_t n) {
	return (uint32_t)hj_bounded(hj_rand(seed, stream, row * 16 + field), n);
}
// append the NUL-terminated src at dst[pos]; returns the new position
__device__ __forceinline__ int put(char *dst, int pos, const char *src) {
	for (int i = 0; src[i]; i++) dst[pos++] = src[i];
	return pos;
}
// days since 1970-01-01 -> "YYYY-MM-DD" (proleptic Gregorian calendar)
__device__ void format_date(int64_t z, char *out) {
	z += 719468;
	const int64_t era = (z >= 0 ? z : z - 146096) / 146097;
	const unsigned doe = (unsigned)(z - era * 146097);
	const unsigned yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365;
	int64_t y = (int64_t)yoe + era * 400;
	const unsigned doy = doe - (365 * yoe + yoe / 4 - yoe / 100);
	const unsigned mp = (5 * doy + 2) / 153;
	const unsigned d = doy - (153 * mp + 2) / 5 + 1;
	const unsigned m = mp < 10 ? mp + 3 : mp - 9;
	y += m <= 2;
	out[0] = (char)('0' + y / 1000);
	out[1] = (char)('0' + y / 100 % 10);
	out[2] = (char)('0' + y / 10 % 10);
	out[3] = (char)('0' + y % 10);
	out[4] = '-';
	out[5] = (char)('0' + m / 10);
	out[6] = (char)('0' + m % 10);
	out[7] = '-';
	out[8] = (char)('0' + d / 10);
	out[9] = (char)('0' + d % 10);
}

__global__ __launch_bounds__(256) void k_gen_part(uint64_t n, uint64_t seed, hj_tpch_part_t p) {
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
		char t[25] = {}, b[10] = {}, c[10] = {};
		int k = put(t, 0, k_type1[pick(seed, TPCH_STREAM_PART, i, 0, 6)]);
		t[k++] = ' ';
		k = put(t, k, k_type2[pick(seed, TPCH_STREAM_PART, i, 1, 5)]);
		t[k++] = ' ';
		put(t, k, k_type3[pick(seed, TPCH_STREAM_PART, i, 2, 5)]);
		put(b, 0, "Brand#");
		b[6] = (char)('1' + pick(seed, TPCH_STREAM_PART, i, 3, 5));
		b[7] = (char)('1' + pick(seed, TPCH_STREAM_PART, i, 4, 5));
		k = put(c, 0, k_cont1[pick(seed, TPCH_STREAM_PART, i, 5, 5)]);
		c[k++] = ' ';
		put(c, k, k_cont2[pick(seed, TPCH_STREAM_PART, i, 6, 8)]);
		p.partkey[i] = i + 1;
		for (int j = 0; j < 25; j++) p.type[i * 25 + j] = t[j];
		for (int j = 0; j < 10; j++) p.brand[i * 10 + j] = b[j];
		for (int j = 0; j < 10; j++) p.container[i * 10 + j] = c[j];
		p.size[i] = 1 + (int32_t)pick(seed, TPCH_STREAM_PART, i, 7, 50);
	}
}

// p_retailprice in cents (TPC-H 4.2.3): 90000 + (key / 10) mod 20001 + 100 * (key mod 1000)
__device__ __forceinline__ int64_t retail_cents(uint64_t key) {
	return 90000 + (int64_t)((key / 10) % 20001) + 100 * (int64_t)(key % 1000);
}

__global__ __launch_bounds__(256) void k_gen_lineitem(uint64_t n, uint64_t nparts, uint64_t seed,
                                                      hj_tpch_lineitem_t l) {
	// o_orderdate uniform over [1992-01-01, 1998-12-31 - 151 days] (day 8035 is
	// 1992-01-01), l_shipdate 1..121 days later
	const int64_t start = 8035;
	const uint32_t span = 2557 - 151;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
		const uint64_t key = 1 + hj_bounded(hj_rand(seed, TPCH_STREAM_LINE, i * 16), nparts);
		const int32_t q = 1 + (int32_t)pick(seed, TPCH_STREAM_LINE, i, 1, 50);
		l.partkey[i] = key;
		l.quantity[i] = q;
		l.extendedprice[i] = q * retail_cents(key);
		l.discount[i] = (int32_t)pick(seed, TPCH_STREAM_LINE, i, 2, 11);
		char d[10];
		format_date(start + pick(seed, TPCH_STREAM_LINE, i, 3, span) + 1 + pick(seed, TPCH_STREAM_LINE, i, 4, 121), d);
		for (int j = 0; j < 10; j++) l.shipdate[i * 10 + j] = d[j];
		char s[25] = {}, m[10] = {};
		put(s, 0, k_instr[pick(seed, TPCH_STREAM_LINE, i, 5, 4)]);
		put(m, 0, k_mode[pick(seed, TPCH_STREAM_LINE, i, 6, 7)]);
		for (int j = 0; j < 25; j++) l.shipinstruct[i * 25 + j] = s[j];
		for (int j = 0; j < 10; j++) l.shipmode[i * 10 + j] = m[j];
	}
}

unsigned tp_grid(hj_ctx_t *ctx, uint64_t n, uint64_t per_block) {
	return (unsigned)std::max<uint64_t>(
	    1, std::min<uint64_t>((n + per_block - 1) / per_block, (uint64_t)ctx->num_cu * 8));
}

// The reference's TPC-H algorithm table (tpch_query.c:93-99) and the join of
// this library each one runs.
struct tpch_algo {
	const char *name, *join;
};
const tpch_algo k_tpch_algos[] = {{"tpch_nphj_sc", "nphj_sc"},
                                  {"tpch_rdx_bc", "phj_radix_bc"},
                                  {"tpch_rdx_hr", "phj_radix_hr"},
                                  {"tpch_rdx_bw_reg_bc", "phj_radix_bw_reg_bc"},
                                  {"tpch_rdx_bw_reg_hr", "phj_radix_bw_reg_hr"}};
const int k_n_tpch_algos = (int)(sizeof k_tpch_algos / sizeof k_tpch_algos[0]);

bool part_cols_ok(const hj_tpch_part_t *p) {
	return !p->rows || (p->partkey && p->brand && p->type && p->size && p->container);
}
bool lineitem_cols_ok(const hj_tpch_lineitem_t *l) {
	return !l->rows || (l->partkey && l->quantity && l->extendedprice && l->discount && l->shipdate &&
	                    l->shipinstruct && l->shipmode);
}

// ---------------------------------------------------------------- text files
// Fields of one '|'-separated line: empty fields kept (dbgen writes none).
int split_fields(char *line, char **f, size_t *len, int max) {
	int k = 0;
	char *s = line;
	while (k < max) {
		char *e = strchr(s, '|');
		if (!e) break;
		f[k] = s;
		len[k] = (size_t)(e - s);
		k++;
		s = e + 1;
	}
	return k;
}
// "123.45" -> 12345: a decimal with at most two fraction digits, exactly
bool parse_hundredths(const char *s, size_t n, int64_t &v) {
	const char *e = s + n;
	bool neg = false;
	if (s < e && *s == '-') neg = true, s++;
	if (s == e) return false;
	int64_t ip = 0, fp = 0;
	int fd = -1;
	for (; s < e; s++) {
		if (*s == '.' && fd < 0) {
			fd = 0;
		} else if (*s >= '0' && *s <= '9') {
			if (fd < 0) {
				if (ip > (int64_t)1e15) return false;
				ip = ip * 10 + (*s - '0');
			} else {
				if (fd >= 2) return false;
				fp = fp * 10 + (*s - '0');
				fd++;
			}
		} else {
			return false;
		}
	}
	for (int i = fd < 0 ? 0 : fd; i < 2; i++) fp *= 10;
	v = (ip * 100 + fp) * (neg ? -1 : 1);
	return true;
}
bool parse_u64(const char *s, size_t n, uint64_t &v) {
	if (!n || n > 19) return false;
	v = 0;
	for (size_t i = 0; i < n; i++) {
		if (s[i] < '0' || s[i] > '9') return false;
		v = v * 10 + (uint64_t)(s[i] - '0');
	}
	return true;
}
void put_field(char *dst, int width, const char *src, size_t n) {
	memset(dst, 0, (size_t)width);
	memcpy(dst, src, std::min<size_t>(n, (size_t)width));
}

// The reference's packed row layouts (schema.h:37-67): part.bin and
// lineitem.bin are arrays of these.
#pragma pack(push, 1)
struct ref_part_row {
	uint64_t p_partkey;
	char p_name[55];
	char p_mfgr[25];
	char p_brand[10];
	char p_type[25];
	int32_t p_size;
	char p_container[10];
	double p_retailprice;
	char p_comment[23];
};
struct ref_lineitem_row {
	uint64_t l_orderkey, l_partkey, l_suppkey, l_linenumber;
	double l_quantity, l_extendedprice, l_discount, l_tax;
	char l_returnflag[1], l_linestatus[1];
	char l_shipdate[10], l_commitdate[10], l_receiptdate[10];
	char l_shipinstruct[25];
	char l_shipmode[10];
	char l_comment[44];
};
#pragma pack(pop)
static_assert(sizeof(ref_part_row) == 168, "part_t is 168 packed bytes");
static_assert(sizeof(ref_lineitem_row) == 175, "lineitem_t is 175 packed bytes");

// strtok(line, "|") tokens, as tpch_tbl2bin.c parses a row: empty fields are
// skipped.  Returns the number of tokens (at most max).
int strtok_fields(char *line, char **tok, int max) {
	int k = 0;
	char *save = nullptr;
	for (char *t = strtok_r(line, "|", &save); t && k < max; t = strtok_r(nullptr, "|", &save)) tok[k++] = t;
	return k;
}

// double money -> exact hundredths; false if it is not one
bool money_to_hundredths(double x, int64_t &v) {
	if (!(fabs(x) < 9e13)) return false;
	const double r = nearbyint(x * 100.0);
	if (fabs(x * 100.0 - r) > 1e-6 * std::max(1.0, fabs(r))) return false;
	v = (int64_t)r;
	return true;
}

}  // namespace

extern "C" {

int hj_tpch_algo_count(void) { return k_n_tpch_algos; }
const char *hj_tpch_algo_name(int idx) { return idx >= 0 && idx < k_n_tpch_algos ? k_tpch_algos[idx].name : nullptr; }

int hj_tpch_cardinality(int sf, uint64_t *part_rows, uint64_t *lineitem_rows) {
	if (!part_rows || !lineitem_rows) return HJ_EINVAL;
	switch (sf) {  // get_cardinality, tpch_utils.c:20-40
	case 1: *part_rows = 200000, *lineitem_rows = 6001215; break;
	case 10: *part_rows = 2000000, *lineitem_rows = 59986052; break;
	case 100: *part_rows = 20000000, *lineitem_rows = 600037902; break;
	default: return HJ_EINVAL;
	}
	return HJ_OK;
}

int hj_tpch_generate(hj_ctx_t *ctx, uint64_t seed, const hj_tpch_part_t *part, const hj_tpch_lineitem_t *li) {
	if (!ctx || !part || !li) return HJ_EINVAL;
	if (!part_cols_ok(part) || !lineitem_cols_ok(li)) return hj_fail(ctx, HJ_EINVAL, "null TPC-H column");
	if (li->rows && !part->rows) return hj_fail(ctx, HJ_EINVAL, "lineitem rows need at least one part row");
	std::lock_guard<std::mutex> lk(ctx->mu);
	HJ_HIP(ctx, hipSetDevice(ctx->device));
	if (part->rows)
		hipLaunchKernelGGL(k_gen_part, dim3(tp_grid(ctx, part->rows, 256)), dim3(256), 0, ctx->stream, part->rows,
		                   seed, *part);
	if (li->rows)
		hipLaunchKernelGGL(k_gen_lineitem, dim3(tp_grid(ctx, li->rows, 256)), dim3(256), 0, ctx->stream, li->rows,
		                   part->rows, seed, *li);
	HJ_HIP(ctx, hipGetLastError());
	HJ_HIP(ctx, hipStreamSynchronize(ctx->stream));
	return HJ_OK;
}

int hj_tpch_query(hj_ctx_t *ctx, int query, const char *algo, const hj_tpch_part_t *part,
                  const hj_tpch_lineitem_t *li, hj_tpch_result_t *out) {
	if (!ctx || !algo || !part || !li || !out) return HJ_EINVAL;
	if (query != 14 && query != 19) return hj_fail(ctx, HJ_EINVAL, "TPC-H query %d: only 14 and 19", query);
	const char *join_algo = nullptr;
	for (int i = 0; i < k_n_tpch_algos; i++)
		if (!strcmp(algo, k_tpch_algos[i].name)) join_algo = k_tpch_algos[i].join;
	if (!join_algo) return hj_fail(ctx, HJ_EINVAL, "unknown TPC-H algo '%s'", algo);
	if (!part_cols_ok(part) || !lineitem_cols_ok(li)) return hj_fail(ctx, HJ_EINVAL, "null TPC-H column");
	memset(out, 0, sizeof *out);
	const uint64_t np = part->rows, nl = li->rows;
	const int pouts = query == 14 ? 2 : 3, louts = query == 14 ? 1 : 3;
	uint4 *po[3], *lo[3];  // the contiguous outputs
	unsigned long long n[8];
	{
		// the pushdown: one selection pass over each table into per-workgroup
		// staging regions, a scan of the workgroup counts, a gather into
		// contiguous outputs of the exact sizes
		std::lock_guard<std::mutex> lk(ctx->mu);
		HJ_HIP(ctx, hipSetDevice(ctx->device));
		uint32_t pblk, lblk;
		const uint64_t pper = query == 14 ? select_per_block<sel_q14_parts>(ctx, np, &pblk)
		                                  : select_per_block<sel_q19_parts>(ctx, np, &pblk);
		const uint64_t lper = query == 14 ? select_per_block<sel_q14_lineitems>(ctx, nl, &lblk)
		                                  : select_per_block<sel_q19_lineitems>(ctx, nl, &lblk);
		int rc;
		if ((rc = hj_reserve(ctx, ctx->tpch[0], (size_t)pouts * std::max<uint64_t>(np, 1) * 4)) ||
		    (rc = hj_reserve(ctx, ctx->tpch[1], (size_t)louts * std::max<uint64_t>(nl, 1) * 4)) ||
		    (rc = hj_reserve(ctx, ctx->tpch[2], ((size_t)pblk * 4 + (size_t)lblk * 4 + 8) * 8)))
			return rc;
		uint32_t *ps = (uint32_t *)ctx->tpch[0].p, *ls = (uint32_t *)ctx->tpch[1].p;
		unsigned long long *pcnt = (unsigned long long *)ctx->tpch[2].p, *lcnt = pcnt + (size_t)pblk * 4,
		                   *tot = lcnt + (size_t)lblk * 4;
		// staging lists a query does not have alias the first (no row
		// selects them)
		uint32_t *pst[3] = {ps, ps + np, query == 14 ? ps : ps + 2 * np};
		uint32_t *lst[3] = {ls, query == 14 ? ls : ls + nl, query == 14 ? ls : ls + 2 * nl};
		const sel_q14_parts q14p{part->partkey, part->type};
		const sel_q14_lineitems q14l{li->partkey, li->extendedprice, li->discount, li->shipdate};
		const sel_q19_parts q19p{part->partkey, part->brand, part->size, part->container};
		const sel_q19_lineitems q19l{li->partkey, li->quantity, li->extendedprice, li->discount, li->shipinstruct,
		                             li->shipmode};
		const int e0 = hj_ctx::EV_PTR;
		HJ_HIP(ctx, hipEventRecord(ctx->ev[e0], ctx->stream));
		if (query == 14) {
			launch_select(ctx, np, pper, pblk, q14p, pst, pcnt);
			launch_select(ctx, nl, lper, lblk, q14l, lst, lcnt);
		} else {
			launch_select(ctx, np, pper, pblk, q19p, pst, pcnt);
			launch_select(ctx, nl, lper, lblk, q19l, lst, lcnt);
		}
		hipLaunchKernelGGL(k_select_scan, dim3(1), dim3(256), 0, ctx->stream, pcnt, pblk, tot);
		hipLaunchKernelGGL(k_select_scan, dim3(1), dim3(256), 0, ctx->stream, lcnt, lblk, tot + 4);
		HJ_HIP(ctx, hipGetLastError());
		HJ_HIP(ctx, hipMemcpyAsync(n, tot, sizeof n, hipMemcpyDeviceToHost, ctx->stream));
		HJ_HIP(ctx, hipStreamSynchronize(ctx->stream));
		const uint64_t pt = n[0] + n[1] + n[2], lt = n[4] + n[5] + n[6];
		if ((rc = hj_reserve(ctx, ctx->tpch[3], std::max<uint64_t>(pt, 1) * 16)) ||
		    (rc = hj_reserve(ctx, ctx->tpch[4], std::max<uint64_t>(lt, 1) * 16)))
			return rc;
		po[0] = (uint4 *)ctx->tpch[3].p, po[1] = po[0] + n[0], po[2] = po[1] + n[1];
		lo[0] = (uint4 *)ctx->tpch[4].p, lo[1] = lo[0] + n[4], lo[2] = lo[1] + n[5];
		if (query == 14) {
			launch_gather(ctx, pouts, pper, pblk, q14p, pst, po, pcnt, tot);
			launch_gather(ctx, louts, lper, lblk, q14l, lst, lo, lcnt, tot + 4);
		} else {
			launch_gather(ctx, pouts, pper, pblk, q19p, pst, po, pcnt, tot);
			launch_gather(ctx, louts, lper, lblk, q19l, lst, lo, lcnt, tot + 4);
		}
		HJ_HIP(ctx, hipGetLastError());
		HJ_HIP(ctx, hipEventRecord(ctx->ev[e0 + 1], ctx->stream));
		HJ_HIP(ctx, hipStreamSynchronize(ctx->stream));
		out->pushdown_time = hj_event_seconds(ctx, e0, e0 + 1);
	}
	out->part_selected = n[3];
	out->lineitem_selected = n[7];
	out->part_tuples = n[0] + n[1] + n[2];
	out->lineitem_tuples = n[4] + n[5] + n[6];
	// the joins: R = selected parts (payload 0), S = selected lineitems
	// (payload = revenue), so check_sum is the revenue over the matching pairs
	auto join = [&](const uint4 *r, uint64_t r_n, const uint4 *s, uint64_t s_n, hj_result_t &res) -> int {
		hj_relation_t R{(void *)r, r_n, 16, HJ_LOC_DEVICE}, S{(void *)s, s_n, 16, HJ_LOC_DEVICE};
		hj_params_t prm{};
		out->joins++;
		const int rc = hj_join(ctx, join_algo, &R, &S, &prm, &res);
		out->buildpart_time += res.buildpart_time;
		out->probejoin_time += res.probejoin_time;
		return rc;
	};
	hj_result_t r;
	int rc;
	if (query == 14) {
		if ((rc = join(po[0], n[0], lo[0], n[4], r))) return rc;
		out->matched_cnt = r.matched_cnt;
		out->denominator_e4 = r.check_sum;
		if ((rc = join(po[1], n[1], lo[0], n[4], r))) return rc;
		out->numerator_e4 = r.check_sum;
		out->numerator = (double)out->numerator_e4 * 1e-4;
		out->denominator = (double)out->denominator_e4 * 1e-4;
		out->promo_revenue = out->denominator_e4 ? 100.0 * (double)out->numerator_e4 / (double)out->denominator_e4 : 0.0;
	} else {
		for (int g = 0; g < 3; g++) {
			if ((rc = join(po[g], n[g], lo[g], n[4 + g], r))) return rc;
			out->matched_cnt += r.matched_cnt;
			out->revenue_e4 += r.check_sum;
		}
		out->revenue = (double)out->revenue_e4 * 1e-4;
	}
	out->total_time = out->pushdown_time + out->buildpart_time + out->probejoin_time;
	return HJ_OK;
}

int hj_tpch_tbl_rows(const char *path, uint64_t *rows) {
	if (!path || !rows) return HJ_EINVAL;
	FILE *f = fopen(path, "rb");
	if (!f) return HJ_EINVAL;
	std::vector<char> buf(1 << 20);
	uint64_t n = 0;
	char last = '\n';
	size_t got;
	while ((got = fread(buf.data(), 1, buf.size(), f)) > 0) {
		for (size_t i = 0; i < got; i++) n += buf[i] == '\n';
		last = buf[got - 1];
	}
	fclose(f);
	*rows = n + (last != '\n');
	return HJ_OK;
}

// dbgen's part.tbl: p_partkey|p_name|p_mfgr|p_brand|p_type|p_size|p_container|
// p_retailprice|p_comment|
int hj_tpch_tbl_read_part(const char *path, const hj_tpch_part_t *p) {
	if (!path || !p || !part_cols_ok(p)) return HJ_EINVAL;
	FILE *f = fopen(path, "rb");
	if (!f) return HJ_EINVAL;
	std::vector<char> line(1 << 16);
	uint64_t i = 0;
	int rc = HJ_OK;
	while (rc == HJ_OK && fgets(line.data(), (int)line.size(), f)) {
		if (line[0] == '\n' || line[0] == 0) continue;
		char *fld[9];
		size_t len[9];
		uint64_t key;
		int64_t size;
		if (i >= p->rows || split_fields(line.data(), fld, len, 9) < 7 || !parse_u64(fld[0], len[0], key) ||
		    !parse_hundredths(fld[5], len[5], size) || size % 100) {
			rc = HJ_EINVAL;
			break;
		}
		p->partkey[i] = key;
		put_field(p->brand + i * 10, 10, fld[3], len[3]);
		put_field(p->type + i * 25, 25, fld[4], len[4]);
		p->size[i] = (int32_t)(size / 100);
		put_field(p->container + i * 10, 10, fld[6], len[6]);
		i++;
	}
	fclose(f);
	return rc == HJ_OK && i != p->rows ? HJ_EINVAL : rc;
}

// dbgen's lineitem.tbl: l_orderkey|l_partkey|l_suppkey|l_linenumber|
// l_quantity|l_extendedprice|l_discount|l_tax|l_returnflag|l_linestatus|
// l_shipdate|l_commitdate|l_receiptdate|l_shipinstruct|l_shipmode|l_comment|
int hj_tpch_tbl_read_lineitem(const char *path, const hj_tpch_lineitem_t *l) {
	if (!path || !l || !lineitem_cols_ok(l)) return HJ_EINVAL;
	FILE *f = fopen(path, "rb");
	if (!f) return HJ_EINVAL;
	std::vector<char> line(1 << 16);
	uint64_t i = 0;
	int rc = HJ_OK;
	while (rc == HJ_OK && fgets(line.data(), (int)line.size(), f)) {
		if (line[0] == '\n' || line[0] == 0) continue;
		char *fld[16];
		size_t len[16];
		uint64_t key;
		int64_t qty, price, disc;
		if (i >= l->rows || split_fields(line.data(), fld, len, 16) < 15 || !parse_u64(fld[1], len[1], key) ||
		    !parse_hundredths(fld[4], len[4], qty) || qty % 100 || !parse_hundredths(fld[5], len[5], price) ||
		    !parse_hundredths(fld[6], len[6], disc) || disc < 0 || disc > 100 || len[10] != 10) {
			rc = HJ_EINVAL;
			break;
		}
		l->partkey[i] = key;
		l->quantity[i] = (int32_t)(qty / 100);
		l->extendedprice[i] = price;
		l->discount[i] = (int32_t)disc;
		memcpy(l->shipdate + i * 10, fld[10], 10);
		put_field(l->shipinstruct + i * 25, 25, fld[13], len[13]);
		put_field(l->shipmode + i * 10, 10, fld[14], len[14]);
		i++;
	}
	fclose(f);
	return rc == HJ_OK && i != l->rows ? HJ_EINVAL : rc;
}

// tbl2bin_part / tbl2bin_lineitem (tpch_tbl2bin.c:23-150): a .tbl file into
// the reference's packed binary rows, converted as it converts them (strtok
// fields, strncpy'd strings, atol / atoi / atof numbers).
int hj_tpch_tbl2bin(const char *table, const char *tbl_path, const char *bin_path, uint64_t *rows) {
	if (!table || !tbl_path || !bin_path) return HJ_EINVAL;
	const bool is_part = !strcmp(table, "part");
	if (!is_part && strcmp(table, "lineitem")) return HJ_EINVAL;
	FILE *in = fopen(tbl_path, "rb");
	if (!in) return HJ_EINVAL;
	FILE *out = fopen(bin_path, "wb");
	if (!out) {
		fclose(in);
		return HJ_EINVAL;
	}
	std::vector<char> line(1 << 16);
	uint64_t n = 0;
	int rc = HJ_OK;
	while (rc == HJ_OK && fgets(line.data(), (int)line.size(), in)) {
		char *t[16];
		const int k = strtok_fields(line.data(), t, 16);
		if (is_part) {
			ref_part_row r;
			memset(&r, 0, sizeof r);
			if (k < 9) {
				rc = HJ_EINVAL;
				break;
			}
			r.p_partkey = (uint64_t)atol(t[0]);
			strncpy(r.p_name, t[1], 55);
			strncpy(r.p_mfgr, t[2], 25);
			strncpy(r.p_brand, t[3], 10);
			strncpy(r.p_type, t[4], 25);
			r.p_size = atoi(t[5]);
			strncpy(r.p_container, t[6], 10);
			r.p_retailprice = atof(t[7]);
			strncpy(r.p_comment, t[8], 23);
			if (fwrite(&r, sizeof r, 1, out) != 1) rc = HJ_EINVAL;
		} else {
			ref_lineitem_row r;
			memset(&r, 0, sizeof r);
			if (k < 16) {
				rc = HJ_EINVAL;
				break;
			}
			r.l_orderkey = (uint64_t)atol(t[0]);
			r.l_partkey = (uint64_t)atol(t[1]);
			r.l_suppkey = (uint64_t)atol(t[2]);
			r.l_linenumber = (uint64_t)atoi(t[3]);
			r.l_quantity = atof(t[4]);
			r.l_extendedprice = atof(t[5]);
			r.l_discount = atof(t[6]);
			r.l_tax = atof(t[7]);
			strncpy(r.l_returnflag, t[8], 1);
			strncpy(r.l_linestatus, t[9], 1);
			strncpy(r.l_shipdate, t[10], 10);
			strncpy(r.l_commitdate, t[11], 10);
			strncpy(r.l_receiptdate, t[12], 10);
			strncpy(r.l_shipinstruct, t[13], 25);
			strncpy(r.l_shipmode, t[14], 10);
			strncpy(r.l_comment, t[15], 44);
			if (fwrite(&r, sizeof r, 1, out) != 1) rc = HJ_EINVAL;
		}
		n++;
	}
	fclose(in);
	if (fclose(out) != 0 && rc == HJ_OK) rc = HJ_EINVAL;
	if (rows) *rows = n;
	return rc;
}

// The reference's packed part.bin / lineitem.bin rows into columns: money
// doubles back to exact cents and hundredths (HJ_EINVAL if one is not).
int hj_tpch_bin_read_part(const char *path, const hj_tpch_part_t *p) {
	if (!path || !p || !part_cols_ok(p)) return HJ_EINVAL;
	FILE *f = fopen(path, "rb");
	if (!f) return HJ_EINVAL;
	std::vector<ref_part_row> buf(4096);
	uint64_t i = 0;
	int rc = HJ_OK;
	size_t got;
	while (rc == HJ_OK && (got = fread(buf.data(), sizeof(ref_part_row), buf.size(), f)) > 0) {
		for (size_t j = 0; j < got; j++, i++) {
			if (i >= p->rows) {
				rc = HJ_EINVAL;
				break;
			}
			const ref_part_row &r = buf[j];
			p->partkey[i] = r.p_partkey;
			memcpy(p->brand + i * 10, r.p_brand, 10);
			memcpy(p->type + i * 25, r.p_type, 25);
			p->size[i] = r.p_size;
			memcpy(p->container + i * 10, r.p_container, 10);
		}
	}
	if (rc == HJ_OK && (ferror(f) || i != p->rows)) rc = HJ_EINVAL;
	fclose(f);
	return rc;
}

int hj_tpch_bin_read_lineitem(const char *path, const hj_tpch_lineitem_t *l) {
	if (!path || !l || !lineitem_cols_ok(l)) return HJ_EINVAL;
	FILE *f = fopen(path, "rb");
	if (!f) return HJ_EINVAL;
	std::vector<ref_lineitem_row> buf(4096);
	uint64_t i = 0;
	int rc = HJ_OK;
	size_t got;
	while (rc == HJ_OK && (got = fread(buf.data(), sizeof(ref_lineitem_row), buf.size(), f)) > 0) {
		for (size_t j = 0; j < got; j++, i++) {
			const ref_lineitem_row &r = buf[j];
			int64_t q, price, disc;
			if (i >= l->rows || !money_to_hundredths(r.l_quantity, q) || q % 100 ||
			    !money_to_hundredths(r.l_extendedprice, price) || !money_to_hundredths(r.l_discount, disc) ||
			    disc < 0 || disc > 100) {
				rc = HJ_EINVAL;
				break;
			}
			l->partkey[i] = r.l_partkey;
			l->quantity[i] = (int32_t)(q / 100);
			l->extendedprice[i] = price;
			l->discount[i] = (int32_t)disc;
			memcpy(l->shipdate + i * 10, r.l_shipdate, 10);
			memcpy(l->shipinstruct + i * 25, r.l_shipinstruct, 25);
			memcpy(l->shipmode + i * 10, r.l_shipmode, 10);
		}
	}
	if (rc == HJ_OK && (ferror(f) || i != l->rows)) rc = HJ_EINVAL;
	fclose(f);
	return rc;
}

}  // extern "C"
