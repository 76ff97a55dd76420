// hj_internal.h -- library-private declarations shared by the HIP sources.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string>
#include <mutex>

#include "hjmi.h"

// Kernel ids for per-kernel timing (hj_kernel_name maps them to text).
enum hj_kid : uint32_t {
	HJ_K_FILL = 0,
	HJ_K_LP_BUILD,
	HJ_K_LP_PROBE,
	HJ_K_LP_FIXUP,
	HJ_K_SC_BUILD,
	HJ_K_SC_PROBE,
	HJ_K_HIST,
	HJ_K_SCAN,
	HJ_K_SCATTER,
	HJ_K_PLAN,
	HJ_K_JOIN,
	HJ_K_SORTCHUNK,
	HJ_K_TILESORT,
	HJ_K_GATHERSORT,
	HJ_K_KEYRID,
	HJ_K_RETRIEVE,
	HJ_K_COUNT
};

// One growable device buffer of the context's scratch arena.
struct hj_buf {
	void *p = nullptr;
	size_t cap = 0;
};

struct hj_ctx {
	int device = 0;
	hipStream_t stream = nullptr;
	hipDeviceProp_t prop{};
	int num_cu = 0;
	std::mutex mu;
	std::string last_error;

	// scratch arena: grown on demand, kept across calls so repeated joins do
	// no allocation inside their timed region
	hj_buf table;      // NPHJ tables
	hj_buf pool;       // NPHJ-SC overflow nodes
	hj_buf aux;        // NPHJ-LP duplicate fix-up indices
	hj_buf part[2];    // radix partition outputs (pass 1, pass 2)
	hj_buf hist;       // radix histogram matrix / scan
	hj_buf meta;       // segments, chunk maps, bounds, work lists
	hj_buf scan_tmp;   // scan partials
	hj_buf stats;      // device counters and results
	void *hstats = nullptr;  // pinned host copy of the counters (hj_stats_read)
	hj_buf hostcopy[2];// device copies of host-resident relations

	// late materialisation: while set, joins on 16-byte tuples also append
	// their (build payload, probe payload) pairs here (emit_cap slots at
	// most, claimed in per-wave chunks of emit_chunk slots)
	uint4 *emit_pairs = nullptr;
	uint64_t emit_cap = 0;
	unsigned emit_chunk = 0;
	// two-pass radix joins of 16-byte tuples: pass B may write 8-byte
	// {key, row_id} tuples when every key and row_id fits 32 bits
	bool narrow_ok = false;
	// ... and pass A may store a tile whose row ids are its positions + 1 as
	// 4-byte {key >> first digit, index in the tile} words (the row id is
	// restored from the position in pass B)
	bool dense_ok = false;
	uint8_t *tile_narrow = nullptr;  // per pass-A tile: 0 16-byte, 1 8-byte, 2 4-byte tuples
	hj_buf pairs;      // the pair buffer of late-materialising joins
	hj_buf keyrid[2];  // {key, row_id} copies of wide R and S
	hj_buf tpch[5];    // TPC-H: staged part / lineitem tuples, selection counts, contiguous outputs

	// no-allocation joins (hj_join_noalloc): while `fixed` is set, the NPHJ
	// runners take their counters (the first 256 bytes), table, overflow pool
	// and fix-up indices from this device region instead of the arena
	uint8_t *fixed = nullptr;
	size_t fixed_cap = 0, fixed_off = 0;

	// size-class autotuner: radix bits / passes per log2 |R| (0xff = unset)
	uint8_t tune_bits[40];
	uint8_t tune_passes[40];
	uint8_t tune_pass1[40];  // first-pass bits of a two-pass class (0 = even split)
	// the fastest one-pass configuration per class (batches and serving
	// partition in one pass): radix bits (0xff = unset) and the scatter's
	// tile shape / line (part_variant bits 0-6)
	uint8_t tune_bits1[40];
	uint8_t tune_var1[40];
	// gathered batches: radix bits per class (0xff = the 6144-tuple rule)
	uint8_t tune_gbits[40];
	bool tuned = false;
	// the same tuned on Zipf-skewed probe sides (hj_autotune_zipf), used by
	// joins with HJ_F_SKEW_HANDLING: bits / passes (0xff = unset), first-pass
	// bits, and the probe-range scale (part_variant bits 20-21)
	uint8_t tune_bits_z[40];
	uint8_t tune_passes_z[40];
	uint8_t tune_pass1_z[40];
	uint8_t tune_pscale_z[40];
	bool tuned_z = false;

	// gathered launches: kernel attributes set, LDS table capacities and
	// resident grids (pass A per first-pass width), queried once
	bool g_ready = false;
	uint32_t g_capn = 0, g_capw = 0;
	size_t g_ldsn = 0, g_ldsw = 0;
	unsigned g_gn = 1, g_gw = 1, g_gA[11] = {};

	// timing
	// ev[0..3] phases of a join, ev[4..] kernel pairs, then three for the
	// outer phases of a late-materialising join and two for the head copy of
	// a wide radix join
	static const int NEV = 2 * HJ_MAX_KTIMES + 9;
	static const int EV_PTR = 2 * HJ_MAX_KTIMES + 4;
	static const int EV_HEAD = 2 * HJ_MAX_KTIMES + 7;  // head copy of a wide radix join
	hipEvent_t ev[NEV] = {};
	bool kernel_times = false;
	int n_kt = 0;
	uint32_t kt_id[HJ_MAX_KTIMES] = {};
	// region i runs from event kt_s[i] to event kt_e[i] (slots of ev[]); a
	// boundary between two regions is one event
	int kt_s[HJ_MAX_KTIMES] = {}, kt_e[HJ_MAX_KTIMES] = {};
	bool kt_open = false;
};

// Set ctx->last_error and return `code`.
int hj_fail(hj_ctx_t *ctx, int code, const char *fmt, ...);
// Ensure `b` holds at least `bytes`; returns HJ_OK or HJ_ENOMEM.
int hj_reserve(hj_ctx_t *ctx, hj_buf &b, size_t bytes);
// Device scratch of `bytes` for the role of `b`: `b` itself, grown on demand,
// or, during a no-allocation join, the next 256-byte-aligned piece of the
// fixed region (HJ_EINVAL if it does not fit; nothing is allocated).
int hj_scratch(hj_ctx_t *ctx, hj_buf &b, size_t bytes, void **p);
// Resolve the no-allocation region of the current device (hj_noalloc.hip);
// called once per context at creation, with the context's device current.
int hj_noalloc_setup(hj_ctx_t *ctx);

#define HJ_HIP(ctx, call)                                                                   \
	do {                                                                                    \
		hipError_t e_ = (call);                                                             \
		if (e_ != hipSuccess)                                                               \
			return hj_fail((ctx), HJ_EHIP, "%s failed: %s (%s:%d)", #call,                  \
			               hipGetErrorString(e_), __FILE__, __LINE__);                       \
	} while (0)

// Per-kernel timing: bracket a launch with kt_begin/kt_end.  kt_next ends
// the open region and opens the next one with a single event (an event
// record costs a few microseconds of device time); kt_open_at / kt_close_at
// take an event of ev[] the caller has just recorded as the region's start
// or end.  kt_ms(i): region i's device milliseconds.
void hj_kt_begin(hj_ctx_t *ctx, uint32_t kid);
void hj_kt_end(hj_ctx_t *ctx);
void hj_kt_next(hj_ctx_t *ctx, uint32_t kid);
void hj_kt_open_at(hj_ctx_t *ctx, uint32_t kid, int ev_index);
void hj_kt_close_at(hj_ctx_t *ctx, int ev_index);
double hj_kt_ms(hj_ctx_t *ctx, int i);

// Device counters block (lives in ctx->stats, or at the start of the fixed
// region during a no-allocation join: see hj_stats_dev).
struct hj_stats {
	unsigned long long matched;    // match count
	unsigned long long checksum;   // payload checksum mod 2^64
	unsigned int flags;            // HJ_DF_* device-detected conditions
	unsigned int pad0;
	unsigned long long pool_next;  // NPHJ-SC overflow node cursor
	unsigned long long work_next;  // persistent-kernel work cursor
	unsigned long long sent_cnt;   // LP: R tuples whose key is the empty marker
	unsigned long long sent_min;   // LP: smallest R row index with that key
	unsigned long long sent_sum;   // payload of that row
	unsigned long long total_work; // work items produced by the planner
	unsigned long long total_chunks;
	unsigned long long pair_next;  // late materialisation: row-id pairs emitted
	unsigned long long retrieved;  // late materialisation: retrieve checksum
	unsigned long long retrieved_ok; // pairs whose row ids lie inside R and S
	unsigned long long narrow_tuples; // tuples of pass-A tiles stored as 8 bytes
	unsigned long long dense_tuples; // tuples of pass-A tiles stored as 4-byte positional words
	unsigned long long pad1;
	unsigned long long xq[8];      // two-pass join: work cursor of each XCD's queue
};

#define HJ_DF_DUPKEY 1u       // duplicate build key met in an LP table
#define HJ_DF_POOL_FULL 2u    // SC overflow pool exhausted
#define HJ_DF_OVERSIZE 4u     // LP partition larger than its LDS table
#define HJ_DF_TABLE_FULL 8u   // LP global table full on insert
#define HJ_DF_WIDE 16u        // a 16-byte tuple whose key or row_id exceeds 32 bits

// ---- launchers (each returns HJ_OK or an error code) -----------------------
int hj_fill16(hj_ctx_t *ctx, void *p, uint64_t n16, uint4 v);
static inline hj_stats *hj_stats_dev(hj_ctx_t *ctx) {
	return ctx->fixed ? (hj_stats *)ctx->fixed : (hj_stats *)ctx->stats.p;
}
int hj_stats_reset(hj_ctx_t *ctx);
// copy the device counters to *h (synchronises the stream)
int hj_stats_read(hj_ctx_t *ctx, hj_stats *h);
int hj_launch_generate(hj_ctx_t *ctx, const hj_gen_spec_t *spec, void *r, void *s);

int hj_run_nphj_lp(hj_ctx_t *ctx, const hj_relation_t *R, const hj_relation_t *S,
                   const hj_params_t *prm, hj_result_t *out);
// first-match join with nphj_lp's semantics and a multiplicative hash (the
// exact path of the radix, batch and serving lp joins)
int hj_run_lp_exact(hj_ctx_t *ctx, const hj_relation_t *R, const hj_relation_t *S, hj_result_t *out);
int hj_run_nphj_sc(hj_ctx_t *ctx, const hj_relation_t *R, const hj_relation_t *S,
                   const hj_params_t *prm, hj_result_t *out);
// tk = per-partition table kind (HJ_TK_* of hj_tables.h: 0 bc, 1 lp, 2 sc,
// 3 hr, 4 hro)
int hj_run_radix(hj_ctx_t *ctx, const hj_relation_t *R, const hj_relation_t *S,
                 const hj_params_t *prm, int tk, hj_result_t *out);
typedef int (*hj_algo_fn)(hj_ctx_t *, const hj_relation_t *, const hj_relation_t *, const hj_params_t *,
                          hj_result_t *);
// heads (first 16 bytes) of n tuples of wv 16-byte words into a dense
// 16-byte relation, on the context's stream
void hj_launch_keyrid(hj_ctx_t *ctx, const void *in, uint64_t n, uint32_t wv, void *out);
// late materialisation around `inner`: join the {key, row_id} words, emit
// row-id pairs, then retrieve the checksum through them
int hj_run_ptr(hj_ctx_t *ctx, const hj_relation_t *R, const hj_relation_t *S, const hj_params_t *prm,
               hj_algo_fn inner, hj_result_t *out);
// radix configuration for |R|: the tuned class if the context has one, else
// the default rule
void hj_radix_config(hj_ctx_t *ctx, uint64_t nr, uint32_t *bits, uint32_t *passes, uint32_t *pass1 = nullptr);
// one-pass radix bits for |R| (batches, serving): the tuned class's (the
// default rule's when untuned, or when tuning left a build side above 2048
// tuples unpartitioned), at most 8; *variant gets the class's scatter shape
uint32_t hj_radix_bits_onepass(hj_ctx_t *ctx, uint64_t nr, uint32_t *variant = nullptr);
extern "C" bool hj_algo_is_lp(const char *algo, bool *radix);

// Time between two recorded events of ctx->ev, in seconds.
double hj_event_seconds(hj_ctx_t *ctx, int a, int b);

static inline uint64_t hj_next_pow2(uint64_t x) {
	uint64_t p = 1;
	while (p < x) p <<= 1;
	return p;
}
static inline uint32_t hj_log2_ceil(uint64_t x) {
	uint32_t b = 0;
	while ((1ULL << b) < x) b++;
	return b;
}
