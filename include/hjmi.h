/*
 * hjmi.h -- C ABI of the MI355X equi-join library (libhjmi.so).
 *
 * Drop-in boundary for the reference's join hot path.  The reference selects
 * a join by name from a table of `void algo(const datameta_t, timekeeper_t*)`
 * entries (src/main.c:66-143, src/algo/nphj.h:80-81, src/algo/phj.h:136-142);
 * each entry loads its own relations, runs build/partition and probe/join
 * phases on pthreads, and prints `matched_cnt` and `check_sum`
 * (src/algo/nphj.c:881, src/algo/phj.c:2206).  This library keeps that shape
 * -- one entry point, algorithm chosen by the same names, two relations in,
 * match count + 64-bit checksum + phase times out -- but takes relations by
 * pointer instead of by file, returns status codes instead of calling exit(),
 * and keeps all state in a caller-owned context instead of file-statics.
 *
 * Only plain C types cross this boundary.  All functions are thread-safe per
 * context; a context serialises its own calls.  Nothing here calls exit().
 */
#ifndef HJMI_H
#define HJMI_H

#include <stdint.h>
#include "hjgen.h"

#ifdef __cplusplus
extern "C" {
#endif
/* the library is built with hidden visibility; only this API is exported */
#pragma GCC visibility push(default)

/* ---- status codes (the reference has none: it asserts or exits,
 *      src/main.c:215-217, src/inc/utils.c:459-461) ------------------------ */
#define HJ_OK 0
#define HJ_EINVAL -1   /* bad argument, unknown algorithm name           */
#define HJ_ENOMEM -2   /* device or host allocation failed               */
#define HJ_EHIP -3     /* a HIP runtime call failed                      */
#define HJ_EREJECT -4  /* serving admission rejected the request         */
#define HJ_ENODEV -5   /* no usable GPU                                  */
#define HJ_EBUSY -6    /* request still running (serve poll)             */

/* ---- relation: replaces relation_t {tuples, tuple_num}
 *      (src/inc/store.h:147-150) plus the tuple width that the reference
 *      fixes at compile time (TUPLE_T_SIZE, src/inc/store.h:109-113) ------- */
#define HJ_LOC_DEVICE 0u  /* `tuples` is a device (HBM) pointer            */
#define HJ_LOC_HOST 1u    /* `tuples` is host memory; copied in, untimed  */

typedef struct {
	void *tuples;         /* packed AoS records {u64 key; u64 row_id; u8 rest[]} */
	uint64_t tuple_num;
	uint32_t tuple_size;  /* bytes per tuple: 16, 32, 64, 128, 256 or 512
	                         (batches and serving: up to 256)            */
	uint32_t location;    /* HJ_LOC_* */
} hj_relation_t;

/* ---- tuning: the reference's compile-time knobs (src/algo/params.h:26-35,
 *      CMakeLists.txt:48-375) become run-time values; 0 = library default
 *      chosen by the size-class autotuner ------------------------------------ */
typedef struct {
	uint32_t radix_bits;     /* total radix bits B (NUM_RADIX_BITS), 0 = auto */
	uint32_t passes;         /* 1 or 2 (NUM_PASSES), 0 = auto                 */
	uint32_t lp_slots_log2;  /* nphj_lp slots = 2^x; 0 = next_pow_2(|R|) as
	                            src/algo/nphj.c:1556-1558                      */
	uint32_t flags;          /* HJ_F_*                                        */
	uint32_t part_variant;   /* tuning sweeps: bits 0-3 partition tile shape
	                            for 16-byte tuples (0 default, 1..5 fixed);
	                            bits 4-6 scatter write-combining line:
	                            0 auto, k = 2^(k-1) tuples; bits 8-11 the
	                            first-pass width of two passes (0 auto);
	                            bits 12-13 the first-pass tile of two passes
	                            (0 auto, 1: 2048, 2: 4096 by 1024 threads,
	                            3: 4096 by 512 threads); bits 14-15 the
	                            second pass's workgroups per CU (0: two,
	                            1: all resident, 2: one); bit 16: first-pass
	                            tiles of 8192 16-byte tuples (1024 x 8);
	                            bits 18-19 = 1: second-pass tiles of 4096
	                            16-byte tuples (0: 2048); bits 20-21: probe
	                            tuples per join work item (0: default, 1:
	                            half, 2: a quarter, 3: double); bits 22-25
	                            = v > 0: second-pass tiles ordered in
	                            blocks of 2^(v-1) first digits (0: 8)       */
	uint32_t radix_bits_pass1; /* first-pass bits of two passes
	                            (NUM_RADIX_BITS_PASS1, phj_asymm_radix.h:10-22);
	                            0 = tuned / half of radix_bits                */
	uint32_t reserved[2];
} hj_params_t;

#define HJ_F_KERNEL_TIMES 1u  /* time each phase with HIP events (two-pass joins: one event between
                                 two phases, so the phases tile the join's device time) */
#define HJ_F_NO_NARROW 2u     /* keep 16-byte tuples through the second radix
                                 pass even when keys and row ids fit 32 bits */
#define HJ_F_NO_POSITIONAL 4u /* two radix passes: store every narrow first-
                                 pass tile as 8-byte {key, row_id}, also where
                                 the row ids are the positions + 1            */
#define HJ_F_BATCH_SCATTER 8u /* batches of 16-byte joins: partition with a
                                 histogram and a one-pass scatter instead of
                                 first-pass tiles and gathered joins          */
#define HJ_F_SKEW_HANDLING 16u /* radix joins with radix_bits = 0: take the
                                 configuration hj_autotune_zipf chose for the
                                 size class (radix bits, passes, first-pass
                                 bits, probe-range scale); the reference's
                                 SKEW_HANDLING build flag (params.h:37-42)    */

/* ---- result: matched_cnt/check_sum as printed by the reference plus the
 *      timekeeper_t phase times (src/inc/store.h:74-78) ---------------------- */
#define HJ_MAX_KTIMES 16
typedef struct {
	uint64_t matched_cnt;    /* pairs (r,s) with r.key == s.key (LP: first match) */
	uint64_t check_sum;      /* sum of r.row_id + s.row_id over matches, mod 2^64  */
	double buildpart_time;   /* seconds, device events: build or partition   */
	double probejoin_time;   /* seconds, device events: probe or join        */
	double total_time;
	uint64_t algo_bytes;     /* algorithmic HBM bytes of the algorithm run    */
	uint32_t radix_bits;     /* what actually ran                             */
	uint32_t passes;
	uint32_t fallback;       /* HJ_FB_* bitmask of exact slow paths taken     */
	uint32_t n_ktimes;
	double ktime_ms[HJ_MAX_KTIMES];  /* per-kernel device ms (HJ_F_KERNEL_TIMES) */
	uint32_t ktime_id[HJ_MAX_KTIMES];
	uint64_t narrow_tuples;  /* two-pass radix, 16-byte tuples: tuples that the
	                            first pass stored as 8-byte {key, row_id}     */
	uint32_t narrow;         /* 1: the second pass and the join moved 8-byte
	                            tuples (every key and row id fit 32 bits);
	                            nphj_lp: its table held 8-byte entries       */
	uint32_t reserved;
	uint64_t positional_tuples; /* of those, tuples of first-pass tiles whose
	                            row ids are their relation positions + 1,
	                            stored as 4-byte {key >> first digit, index in
	                            the tile} words (row id restored from the
	                            position by the second pass)                  */
} hj_result_t;

#define HJ_FB_LP_DUPLICATES 1u  /* duplicate build keys under LP: re-run exact */
#define HJ_FB_LP_OVERSIZE 2u    /* partition above LDS capacity under LP        */
#define HJ_FB_SC_REBUILD 4u     /* nphj_sc overflow pool grown and rebuilt      */

typedef struct hj_ctx hj_ctx_t;

/* Context on one GPU: stream, events and a grow-only scratch arena.
 * Replaces the reference's per-run allocations (src/algo/phj.c:1917-1950,
 * src/algo/nphj.c:746-757) and its file-static hashmask/bitskip
 * (src/algo/nphj.c:95-96). */
int hj_ctx_create(int device, hj_ctx_t **out);
int hj_ctx_destroy(hj_ctx_t *ctx);
const char *hj_strerror(int status);
const char *hj_last_error(hj_ctx_t *ctx);   /* detail of the last failure */

/* Algorithm registry, same names as the reference's algos[] table
 * (src/main.c:72-143): nphj_sc, nphj_lp, phj_radix_{bc,lp,sc,hr,hro},
 * phj_radix_bw_reg_{bc,lp,sc,hr,hro}, and the late-materialising
 * ptr_nphj_{sc,lp}, ptr_phj_radix_{bc,sc,lp,hr},
 * ptr_phj_radix_bw_reg_{bc,sc,lp,hr} (src/main.c:87-92, :119-122).  A ptr_
 * join matches the 16-byte {key, row_id} heads, writes one row-id pair per
 * match and returns check_sum = sum of R[r_row_id-1].row_id +
 * S[s_row_id-1].row_id over the pairs (src/inc/utils.c:712-723); row ids
 * must be 1-based positions (otherwise HJ_EINVAL, counting the bad pairs).
 * Index past the end -> NULL (the reference table has no sentinel,
 * src/main.c:206). */
int hj_algo_count(void);
const char *hj_algo_name(int idx);
const char *hj_kernel_name(uint32_t id);

/* One join: the drop-in for `algo(datameta, &timekeeper)`.  R is the build
 * side, S the probe side.  Borrowed read-only; all scratch is the context's.
 * On HJ_OK `out` holds count, checksum and phase times. */
int hj_join(hj_ctx_t *ctx, const char *algo, const hj_relation_t *R, const hj_relation_t *S,
            const hj_params_t *params, hj_result_t *out);

/* One join that allocates nothing.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory while they run (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * algo: nphj_sc or nphj_lp; R and S device-resident (HJ_LOC_DEVICE); params
 * as hj_join (flags: HJ_F_KERNEL_TIMES only); results identical to hj_join.
 * Its counters, table, overflow pool and fix-up indices live in one static
 * device array of HJ_NOALLOC_SCRATCH_BYTES per GPU, held until the process
 * ends: 256 B of counters, then nphj_lp 24 B per table slot
 * (next_pow_2(|R|) slots: |R| <= 2^21) and nphj_sc 32 B per next_pow_2(|R|)
 * plus 64 B per build tuple and 64 KiB (|R| <= 2^19).  A larger join is
 * rejected with HJ_EINVAL before anything runs.  All calls on one GPU share
 * that GPU's array: calls that overlap on one GPU (several contexts or host
 * threads at once) race. */
#define HJ_NOALLOC_SCRATCH_BYTES (64ull << 20)
int hj_join_noalloc(hj_ctx_t *ctx, const char *algo, const hj_relation_t *R, const hj_relation_t *S,
                    const hj_params_t *params, hj_result_t *out);

/* ---- batched joins (no reference counterpart: the reference runs one join
 * per process).  Many independent joins run as one captured HIP graph: one
 * partition launch over all joins (radix bits chosen per join by size),
 * then one persistent join kernel pulling (join, partition, probe tile)
 * items from a device work queue.  Batches of 16-byte tuples sort every
 * relation's tiles by the join's digit in place and join each partition by
 * gathering its runs from those tiles (an lp join that meets a duplicate
 * build key or a partition larger than one LDS table is recomputed by the
 * exact first-match path); bc batches of wider tuples and with
 * HJ_F_BATCH_SCATTER partition with a histogram and a one-pass scatter, and
 * lp batches off the gathered path run every join through the exact path.
 * Results are per join and bit-identical to hj_join on the same relations.
 * Relations are borrowed for the life of the batch and must stay in device
 * memory; their contents may change between runs. */
typedef struct {
	uint64_t matched_cnt;
	uint64_t check_sum;
} hj_join_result_t;

typedef struct {
	double device_ms;        /* graph execution, device events             */
	double fallback_ms;      /* host-driven exact re-runs (LP duplicates)  */
	uint32_t njoins;
	uint32_t n_unpartitioned;/* joins whose build side fits one LDS table  */
	uint32_t n_partitioned;
	uint32_t n_fallback;     /* joins recomputed by the exact LP path      */
	uint64_t work_items;     /* (join, partition, tile) items in the queue */
	uint64_t tuples;         /* sum of |R|+|S|                             */
	uint64_t scratch_bytes;  /* device memory held by the batch            */
	uint64_t algo_bytes;     /* algorithmic HBM bytes of the last run: each
	                            pass reads and writes its tuples at the
	                            layout it chose (4-, 8- or 16-byte), the join
	                            reads them once                             */
	uint32_t gathered;       /* 1: tile-sorted and gathered, 0: scattered  */
	uint32_t n_ktimes;       /* a gathered batch created with
	                            HJ_F_KERNEL_TIMES runs its kernels without
	                            the graph, each between two events:        */
	double ktime_ms[HJ_MAX_KTIMES];  /* per-kernel device ms               */
	uint32_t ktime_id[HJ_MAX_KTIMES];/* hj_kernel_name ids                 */
} hj_batch_info_t;

typedef struct hj_batch hj_batch_t;

/* algo: phj_radix_bc or phj_radix_lp (or their bw_reg names).
 * params: radix_bits/passes must be 0 (chosen per join); flags honoured. */
int hj_batch_create(hj_ctx_t *ctx, const char *algo, uint32_t njoins, const hj_relation_t *R,
                    const hj_relation_t *S, const hj_params_t *params, hj_batch_t **out);
int hj_batch_run(hj_batch_t *batch, hj_join_result_t *results, hj_batch_info_t *info);
int hj_batch_destroy(hj_batch_t *batch);

/* Size-class autotuner: times each candidate (radix bits, passes) on seeded
 * joins of |R| = 2^c, |S| = s_ratio * |R| for c in [min_log2, max_log2] and
 * keeps the fastest per class in the context; batches and hj_join with
 * radix_bits = 0 then use it.  hj_tuned_config reports the choice for |R|. */
int hj_autotune(hj_ctx_t *ctx, uint32_t width, uint32_t s_ratio, uint32_t min_log2, uint32_t max_log2);
int hj_tuned_config(hj_ctx_t *ctx, uint64_t r_num, uint32_t *radix_bits, uint32_t *passes);
/* the tuned first-pass width of a two-pass class (0: the default split) */
int hj_tuned_pass1_bits(hj_ctx_t *ctx, uint64_t r_num, uint32_t *radix_bits_pass1);
/* the radix bits a gathered batch gives a join of |R| = r_num (tuned for
 * 16-byte tuples by hj_autotune, else about 6144 build tuples a partition) */
int hj_tuned_batch_bits(hj_ctx_t *ctx, uint64_t r_num, uint32_t *radix_bits);
/* The autotuner on skewed data: each class's candidates (as hj_autotune)
 * timed on probe keys Zipf-distributed over the build keys with exponent
 * theta, then the probe-range scale (part_variant bits 20-21) for the best;
 * joins with HJ_F_SKEW_HANDLING use these choices.  hj_tuned_skew_config
 * reports them for |R| (probe_scale: the bits 20-21 value). */
int hj_autotune_zipf(hj_ctx_t *ctx, uint32_t width, uint32_t s_ratio, uint32_t min_log2, uint32_t max_log2,
                     double theta);
int hj_tuned_skew_config(hj_ctx_t *ctx, uint64_t r_num, uint32_t *radix_bits, uint32_t *passes,
                         uint32_t *radix_bits_pass1, uint32_t *probe_scale);

/* ---- serving: continuous batching of arriving joins (no reference
 * counterpart).  A scheduler thread owns `slots` HIP streams, each with an
 * HBM arena of hbm_cap_bytes / slots (half the cap when solo lanes exist;
 * the other half bounds the lanes).  Requests queue in arrival order;
 * whenever a slot is free, requests from the head of the queue are admitted
 * while their exact scratch footprint fits that slot's arena (at most
 * max_batch per launch), planned as one batch and launched asynchronously,
 * so batches on different slots overlap on the device.  16-byte servers
 * run their batches gathered (as hj_batch_* does) and batch every request
 * whose build side holds at most 2^20 tuples (2^22 without solo lanes);
 * other bc servers (wider tuples, HJ_SERVE_SCATTER) partition batches with
 * a histogram and a one-pass scatter, and lp requests that do not run
 * gathered are joined by the exact first-match path.  Larger requests (gathered servers) or
 * those with |R| >= 2^solo_min_log2 (scattering ones) go to solo lanes:
 * worker threads with their own context and stream that run the full
 * single-join engine (two radix passes) one request at a time, overlapping
 * with the batches.  Registered build relations are partitioned once and
 * stay resident in HBM; requests naming them only partition their probe
 * side.  A request whose footprint exceeds one arena is rejected
 * (HJ_EREJECT).  Results are bit-identical to hj_join on the same
 * relations.  All relations are borrowed until their request (or, for
 * builds, the server) is done. */
#define HJ_SERVE_SCATTER 1u  /* flags: never gathered batches              */
typedef struct {
	uint64_t hbm_cap_bytes;  /* scratch cap over batch slots and solo lanes */
	uint32_t slots;          /* concurrent batches (streams), 1..8          */
	uint32_t max_batch;      /* joins per launch                            */
	uint32_t solo_lanes;     /* lanes for large joins, 0..4 (0: none)       */
	uint32_t solo_min_log2;  /* scattering servers: |R| >= 2^x runs in a
	                            solo lane (0: 16)                           */
	uint32_t flags;          /* HJ_SERVE_*                                  */
	uint32_t reserved;
} hj_serve_config_t;

typedef struct {
	uint64_t submitted, completed, rejected, batches;
	uint64_t max_batch_joins;      /* largest batch launched               */
	uint32_t max_busy_slots;       /* most batches in flight at once       */
	uint32_t n_builds;
	uint64_t resident_bytes;       /* HBM held by resident build tables    */
	uint64_t arena_bytes;          /* HBM held by the slot arenas          */
	uint64_t peak_admitted_bytes;  /* largest footprint in flight at once  */
	uint64_t solo_completed;       /* requests run by the solo lanes        */
} hj_serve_stats_t;

typedef struct hj_server hj_server_t;

int hj_serve_open(hj_ctx_t *ctx, const char *algo, const hj_serve_config_t *cfg, hj_server_t **out);
/* partition R once and keep it resident; returns its id */
int hj_serve_register_build(hj_server_t *srv, const hj_relation_t *R, uint32_t *build_id);
/* build_id < 0: R is the build side; otherwise R is ignored */
int hj_serve_submit(hj_server_t *srv, int32_t build_id, const hj_relation_t *R, const hj_relation_t *S,
                    uint64_t *req_id);
/* HJ_EBUSY until done; latency = completion - submission, milliseconds */
int hj_serve_poll(hj_server_t *srv, uint64_t req_id, hj_join_result_t *res, double *latency_ms);
int hj_serve_drain(hj_server_t *srv);
int hj_serve_stats(hj_server_t *srv, hj_serve_stats_t *st);
int hj_serve_close(hj_server_t *srv);

/* ---- device memory helpers for hosts without an allocator of their own --- */
int hj_dev_alloc(hj_ctx_t *ctx, uint64_t bytes, void **ptr);
int hj_dev_free(hj_ctx_t *ctx, void *ptr);
int hj_copy_h2d(hj_ctx_t *ctx, void *dst, const void *src, uint64_t bytes);
int hj_copy_d2h(hj_ctx_t *ctx, void *dst, const void *src, uint64_t bytes);
int hj_sync(hj_ctx_t *ctx);
/* Device-to-device copy by a plain streaming kernel on the context's stream
 * (no reference counterpart; used to duplicate relations in HBM and as the
 * known-size copy that calibrates HBM byte counters).  Returns the kernel's
 * device time in *ms when ms is not NULL.  bytes must be a multiple of 16. */
int hj_copy_d2d(hj_ctx_t *ctx, void *dst, const void *src, uint64_t bytes, double *ms);

/* ---- seeded generator on the device (replaces bin/gen, src/gen.c:29-158).
 * Fills R (spec->r_num tuples) and S (spec->s_num tuples) of width
 * spec->width in device memory allocated by the caller.  Byte-identical to
 * the host functions in hjgen.h. */
int hj_generate(hj_ctx_t *ctx, const hj_gen_spec_t *spec, void *r_dev, void *s_dev);

/* ---- the reference's own datasets -----------------------------------------
 * Generator with the reference's draws (src/gen.c:29-377 and
 * src/inc/eth_genzipf.c:30-148 through glibc rand() after srand(seed)):
 * fills host buffers of r_num / s_num tuples.  With exact_order the key and
 * row_id words equal the reference's .bin files byte for byte (its shuffles
 * included; the rest of each tuple is zero); without it the same tuples come
 * unshuffled, which no join result depends on. */
#define HJ_REF_PKFK 0     /* uniform_pkfk_{A,B,C}: gen_pkfk_uniform            */
#define HJ_REF_DUPFK 1    /* uniform_dupfk_{D,E,F}: gen_dup_uniform            */
#define HJ_REF_ZIPF 2     /* zipfian_<1+theta>_A: pk_AR + gen_fk_zipfian_eth   */
#define HJ_REF_PAYLOAD 3  /* payload_pkfk_<bytes>: payload_gen_pkfk_uniform    */
typedef struct {
	uint32_t kind;         /* HJ_REF_*                                      */
	uint32_t tuple_size;   /* bytes, multiple of 16                          */
	uint64_t r_num, s_num; /* pkfk/dupfk/payload: s_num a multiple of r_num,
	                          r_num * tuple_size a multiple of 4096 (whole
	                          blocks, as gen_fk_from_pk copies)              */
	double zipf_factor;    /* zipf: 1 + the configuration's theta            */
	int32_t seed;          /* 12345 in the reference (src/gen.c:144)         */
	uint32_t exact_order;  /* 1: the reference's tuple order too             */
} hj_ref_spec_t;
/* spec of a workload named as the reference's configurations name it
 * (workload_name of the config/data .cfg files: "uniform_dupfk_D",
 * "zipfian_1.250_A", "payload_pkfk_64", ...); exact_order left 0 */
int hj_ref_spec_for(const char *workload_name, uint32_t tuple_size, hj_ref_spec_t *out);
int hj_ref_generate(const hj_ref_spec_t *spec, void *r_host, void *s_host);
/* the first n values of glibc rand() after srand(seed) */
int hj_ref_rand(int32_t seed, uint64_t n, int32_t *out);

/* A data configuration file of the reference (the .cfg files under config/data, read by
 * data_cfg, src/inc/utils.c:236-290). */
typedef struct {
	char r_path_suffix[256], s_path_suffix[256], workload_name[128];
	uint64_t r_tuple_num, s_tuple_num;
	uint32_t r_tuple_size, s_tuple_size;
	double theta, selectivity, density, sparsity;
	int64_t min_key, max_key;
} hj_data_cfg_t;
int hj_data_cfg_read(const char *cfg_path, hj_data_cfg_t *out);
/* Raw tuple files (packed tuple_t[], no header): the reference's data files
 * (read at src/algo/phj.c:1885-1915) into / out of device memory through
 * pinned staging buffers. */
int hj_bin_read(hj_ctx_t *ctx, const char *path, uint64_t tuple_num, uint32_t tuple_size, void *dev);
int hj_bin_write(hj_ctx_t *ctx, const char *path, const void *dev, uint64_t tuple_num, uint32_t tuple_size);

/* ---- TPC-H Q14 and Q19 on the join engine (src/tpch/tpch_query.c) --------
 * The columns the two queries read, one array per column.  Strings are
 * fixed-width, NUL-padded char arrays of the reference's widths
 * (src/tpch/inc/schema.h:37-67): row i's p_type is type[25 * i .. 25 * i + 24].
 * Money is exact: l_extendedprice in cents, l_discount in hundredths,
 * l_quantity in whole units.  Device pointers for hj_tpch_generate and
 * hj_tpch_query, host pointers for the file readers. */
typedef struct {
	uint64_t rows;
	uint64_t *partkey;
	char *brand;       /* 10 bytes a row, "Brand#MN"                    */
	char *type;        /* 25 bytes                                      */
	int32_t *size;
	char *container;   /* 10 bytes                                      */
} hj_tpch_part_t;
typedef struct {
	uint64_t rows;
	uint64_t *partkey;
	int32_t *quantity;
	int64_t *extendedprice;  /* cents                                    */
	int32_t *discount;       /* hundredths (0..100)                      */
	char *shipdate;          /* 10 bytes, "YYYY-MM-DD"                   */
	char *shipinstruct;      /* 25 bytes                                 */
	char *shipmode;          /* 10 bytes                                 */
} hj_tpch_lineitem_t;
/* What tpch_spj prints (src/tpch/tpch_query.c:739-745) plus its phase times.
 * The sums of l_extendedprice * (1 - l_discount) are exact integers in units
 * of 1e-4 (the reference adds doubles); the doubles are those sums / 1e4. */
typedef struct {
	uint64_t matched_cnt;      /* Q14: join pairs; Q19: pairs passing match_filter */
	uint64_t numerator_e4;     /* Q14: PROMO pairs                              */
	uint64_t denominator_e4;   /* Q14: all pairs                                */
	uint64_t revenue_e4;       /* Q19                                           */
	double promo_revenue;      /* Q14: 100 * numerator / denominator             */
	double numerator, denominator, revenue;
	uint64_t part_selected;    /* rows passing part_filter (tpch_utils.c:44-87)  */
	uint64_t lineitem_selected;/* rows passing lineitem_filter (:89-112)         */
	uint64_t part_tuples;      /* {partkey, 0} tuples the part selection wrote
	                              (Q14: all parts + PROMO parts)                 */
	uint64_t lineitem_tuples;  /* {partkey, revenue} tuples the lineitem
	                              selection wrote (Q19: one per group it is in) */
	uint32_t joins;            /* hj_join calls the query made                   */
	uint32_t reserved;
	double pushdown_time, buildpart_time, probejoin_time, total_time;  /* seconds */
} hj_tpch_result_t;
/* The reference's TPC-H algorithm names (tpch_query.c:93-99): tpch_nphj_sc,
 * tpch_rdx_bc, tpch_rdx_hr, tpch_rdx_bw_reg_bc, tpch_rdx_bw_reg_hr. */
int hj_tpch_algo_count(void);
const char *hj_tpch_algo_name(int idx);
/* part / lineitem rows of scale factor 1, 10, 100 (get_cardinality,
 * src/tpch/inc/tpch_utils.c:20-40) */
int hj_tpch_cardinality(int sf, uint64_t *part_rows, uint64_t *lineitem_rows);
/* Seeded TPC-H-like columns in device memory: the value domains of the TPC-H
 * specification (p_partkey = 1..rows, l_partkey uniform over them), drawn by
 * this library's counter-based generator -- not dbgen's streams. */
int hj_tpch_generate(hj_ctx_t *ctx, uint64_t seed, const hj_tpch_part_t *part, const hj_tpch_lineitem_t *lineitem);
/* Q14 or Q19 over device columns: the selections of tpch_utils.c:44-160, the
 * join on partkey by `algo`, the sums of tpch_nphj.c:302-319. */
int hj_tpch_query(hj_ctx_t *ctx, int query, const char *algo, const hj_tpch_part_t *part,
                  const hj_tpch_lineitem_t *lineitem, hj_tpch_result_t *out);
/* dbgen .tbl files ('|'-separated rows) into host columns of exactly `rows`
 * rows; hj_tpch_tbl_rows counts the rows of a file. */
int hj_tpch_tbl_rows(const char *path, uint64_t *rows);
int hj_tpch_tbl_read_part(const char *path, const hj_tpch_part_t *host);
int hj_tpch_tbl_read_lineitem(const char *path, const hj_tpch_lineitem_t *host);
/* tbl2bin (src/tpch/tpch_tbl2bin.c:23-150): table "part" or "lineitem" from a
 * .tbl file into the reference's packed part.bin / lineitem.bin rows (168 and
 * 175 bytes), and those files back into host columns. */
int hj_tpch_tbl2bin(const char *table, const char *tbl_path, const char *bin_path, uint64_t *rows);
int hj_tpch_bin_read_part(const char *path, const hj_tpch_part_t *host);
int hj_tpch_bin_read_lineitem(const char *path, const hj_tpch_lineitem_t *host);

/* ---- device properties used by the autotuner and the benchmark ---------- */
typedef struct {
	int cu_count;
	int lds_per_block;
	int lds_per_cu;
	int warp_size;
	uint64_t hbm_bytes;
	int clock_khz;
	char arch[64];
} hj_device_info_t;
int hj_device_info(hj_ctx_t *ctx, hj_device_info_t *info);

#pragma GCC visibility pop
#ifdef __cplusplus
}
#endif

#endif /* HJMI_H */
