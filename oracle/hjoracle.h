/*
 * hjoracle.h -- TEST INFRASTRUCTURE ONLY.  CPU oracle for the equi-join hot
 * path.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
 * may load it, and only as the checker; the product never links it.
 *
 * Two kinds of functions:
 *   - restatements of the reference's DRAM-mode joins (USE_NVM off), in
 *     their structure: pthreads, static input slices, barriers, per-thread
 *     histograms + cross-thread prefix, 64-byte software write-combining
 *     buffers, a mutex task queue, bc_join / lp_join per partition;
 *   - independent exact checkers: per-key aggregation by sorting, and a
 *     brute-force nested loop for tiny inputs.
 *
 * Semantics shared by every function (the reference's, src/algo/nphj.c:
 * 348-350, src/algo/phj.c:555-557):
 *   matched_cnt = #pairs (r,s) with r.key == s.key
 *   check_sum   = sum of (r.row_id + s.row_id) over those pairs, mod 2^64
 * Linear probing (lp) keeps only the first build tuple of a key in R's input
 * order (src/algo/nphj.c:1162-1165, src/algo/phj.c:655-666).  Departures from
 * the reference, each a reference bug or undefined case:
 *   - a probe miss adds 0 (nphj_lp dereferences NULL there, nphj.c:1253);
 *   - probes are bounded by the table size (the reference spins forever on a
 *     miss in a full table unless TEST_SELECTIVITY is set, phj.c:633-653);
 *   - lp tables mark occupancy separately, so key 0 is an ordinary key
 *     (phj.c:614 uses key 0 as "empty").
 */
#ifndef HJORACLE_H
#define HJORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/hjgen.h"

typedef struct {
	uint64_t matched_cnt;
	uint64_t check_sum;
	double buildpart_time; /* seconds, thread 0 wall clock between barriers */
	double probejoin_time;
	double total_time;
} or_result_t;

/* fill host buffers with the generator of hjgen.h (threads >= 1) */
void or_generate(const hj_gen_spec_t *spec, void *r, void *s, int threads);

/* restatements of the reference joins; width in bytes (16..256) */
int or_nphj_sc(const void *R, uint64_t nr, const void *S, uint64_t ns, uint32_t width, int threads, or_result_t *out);
int or_nphj_lp(const void *R, uint64_t nr, const void *S, uint64_t ns, uint32_t width, int threads, or_result_t *out);
/* kind = per-partition join: 0 bc_join, 1 lp_join, 2 sc_join, 3 hr_join,
 * 4 hro_join (phj.c:496-1037) */
int or_phj_radix(const void *R, uint64_t nr, const void *S, uint64_t ns, uint32_t width, uint32_t radix_bits,
                 uint32_t passes, int kind, int threads, or_result_t *out);

/* asymmetric two passes (src/algo/phj_asymm_radix.c): radix_bits_pass1 bits
 * in the first pass; radix_bits_pass1 >= radix_bits runs one pass */
int or_phj_asymm_radix(const void *R, uint64_t nr, const void *S, uint64_t ns, uint32_t width, uint32_t radix_bits,
                       uint32_t radix_bits_pass1, int kind, int threads, or_result_t *out);

/* stable radix partition of one relation by key bits [shift, shift+bits)
 * (the reference's parallel_radix_partition_optimized, phj.c:257-359, on
 * `threads` slices); hist receives 2^bits partition sizes */
int or_radix_partition(const void *in, uint64_t n, uint32_t width, uint32_t shift, uint32_t bits, int threads,
                       void *out, uint64_t *hist);

/* late materialisation (src/algo/ptr_*.c): count of pairs and the retrieve
 * checksum sum(R[r_row_id-1].row_id + S[s_row_id-1].row_id) of
 * src/inc/utils.c:712-723.  algo 0 nphj_sc, 1 nphj_lp, 2 phj_radix (kind as
 * above).  Returns -1 when a matched row id lies outside its relation. */
int or_ptr_join(int algo, int kind, uint32_t radix_bits, uint32_t passes, const void *R, uint64_t nr,
                const void *S, uint64_t ns, uint32_t width, int threads, or_result_t *out);
/* the literal pipeline for small inputs: emit pairs by nested loop, retrieve */
int or_ptr_bruteforce(const void *R, uint64_t nr, const void *S, uint64_t ns, uint32_t width, int lp,
                      or_result_t *out);
int or_ptr_retrieve(const uint64_t *pairs, uint64_t npairs, const void *R, uint64_t nr, const void *S,
                    uint64_t ns, uint32_t width, uint64_t *checksum);

/* the reference's generator on this C library's srand()/rand() (src/gen.c);
 * kind 0 pkfk, 1 dupfk, 2 zipf (zf = 1 + theta), 3 payload; exact = with the
 * reference's shuffles */
int or_ref_generate(int kind, uint64_t nr, uint64_t ns, uint32_t width, double zf, int seed, int exact, int threads,
                    void *R, void *S);

/* independent checkers */
int or_aggregate(const void *R, uint64_t nr, const void *S, uint64_t ns, uint32_t width, int lp, or_result_t *out);
int or_bruteforce(const void *R, uint64_t nr, const void *S, uint64_t ns, uint32_t width, int lp, or_result_t *out);

#endif
