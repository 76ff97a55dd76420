/*
 * hjgen.h -- seeded relation generator shared by the HIP library (device and
 * host code) and by the CPU test oracle, so both see byte-identical inputs.
 *
 * It replaces the reference's offline generator (src/gen.c:29-331), which
 * writes tuple_t[] files with glibc rand() + Knuth shuffle.  Here every tuple
 * is a pure function of (seed, relation stream, row index):
 *
 *   - counter-based PRNG: SplitMix64 output function keyed per stream;
 *   - permutations of [0, n): a 4-round Feistel network on the next even
 *     power-of-two domain, with cycle walking into [0, n) (a bijection for
 *     every n >= 1, evaluable per element on the GPU);
 *   - Zipf ranks: inverse-CDF lookup in a host-built table of doubles.
 *
 * Tuple layout is the reference's tuple_t (src/inc/store.h:87-96): a packed
 * AoS record of `width` bytes = { u64 key; u64 row_id; u8 values[width-16] }.
 * The join checksum reads row_id (bytes 8..15), called "payload" here.
 *
 * Plain C99; compiled by gcc (oracle, CLI) and hipcc (library).
 */
#ifndef HJGEN_H
#define HJGEN_H

#include <stdint.h>

#if defined(__HIPCC__)
#define HJ_HD __host__ __device__ static inline
#else
#define HJ_HD static inline
#endif

/* ---- workloads ------------------------------------------------------------ */
/* Uniform PK-FK as the C1/C2 configurations state it: R.key is a seeded
 * permutation of 1..|R| with R.payload = key; S.key is drawn uniformly (with
 * replacement) from R's keys; S.payload is the 1-based row index.            */
#define HJ_WL_PKFK_UNIFORM 0
/* Reference-style PK-FK (src/gen.c:199-223): R.key = R.payload = permutation
 * of 1..|R|; S holds |S|/|R| copies of every R key in shuffled order with
 * S.payload = key.  When |R| divides |S| every key appears exactly |S|/|R|
 * times, which gives checksum = |S|*(|R|+1) mod 2^64.                        */
#define HJ_WL_PKFK_REF 1
/* Reference payload-style PK-FK (src/gen.c:226-283): as PKFK_REF for keys,
 * but both payloads are the 1-based row index after shuffling.              */
#define HJ_WL_PAYLOAD_REF 2
/* Skewed, wide: R.key = permutation of 1..|R|, S.key = R key of Zipf rank
 * (rank -> key through a second seeded permutation); payload and the rest of
 * each tuple are seeded random bytes.                                        */
#define HJ_WL_ZIPF 3
/* Keys drawn uniformly from [1, key_range] on both sides (duplicates and
 * misses on both sides); payloads are seeded random.                         */
#define HJ_WL_RANDOM 4

/* relation streams (keep distinct so R and S draws never correlate) */
#define HJ_STREAM_R_PERM 1u
#define HJ_STREAM_S_PERM 2u
#define HJ_STREAM_S_KEY 3u
#define HJ_STREAM_R_PAY 4u
#define HJ_STREAM_S_PAY 5u
#define HJ_STREAM_R_REST 6u
#define HJ_STREAM_S_REST 7u
#define HJ_STREAM_ZIPF_PERM 8u
#define HJ_STREAM_R_KEY 9u

#define HJ_GOLDEN 0x9E3779B97F4A7C15ULL

HJ_HD uint64_t hj_mix64(uint64_t z) {
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
	return z ^ (z >> 31);
}

/* i-th output of a SplitMix64 sequence whose state is keyed by (seed, stream) */
HJ_HD uint64_t hj_rand(uint64_t seed, uint32_t stream, uint64_t i) {
	uint64_t k = hj_mix64(seed + HJ_GOLDEN * ((uint64_t)stream + 1u));
	return hj_mix64(k + HJ_GOLDEN * (i + 1u));
}

/* floor(r * n / 2^64): maps a 64-bit draw to [0, n) */
HJ_HD uint64_t hj_bounded(uint64_t r, uint64_t n) {
#if defined(__HIP_DEVICE_COMPILE__)
	return __umul64hi(r, n);
#else
	return (uint64_t)(((unsigned __int128)r * n) >> 64);
#endif
}

/* ---- permutation of [0, n) ------------------------------------------------ */
typedef struct {
	uint64_t n;
	uint64_t half_mask;
	uint32_t half_bits;
	uint32_t pad_;
	uint64_t round_key[4];
} hj_perm_t;

HJ_HD hj_perm_t hj_perm_make(uint64_t n, uint64_t seed, uint32_t stream) {
	hj_perm_t p;
	uint32_t bits = 0;
	while (bits < 64 && (1ULL << bits) < n) bits++;
	if (bits < 2) bits = 2;
	p.n = n;
	p.half_bits = (bits + 1) / 2;
	p.half_mask = (p.half_bits >= 64) ? ~0ULL : ((1ULL << p.half_bits) - 1);
	p.pad_ = 0;
	for (int r = 0; r < 4; r++) p.round_key[r] = hj_rand(seed, stream, 0x100000000ULL + (uint64_t)r);
	return p;
}

HJ_HD uint64_t hj_perm_feistel(const hj_perm_t *p, uint64_t x) {
	uint64_t l = x >> p->half_bits, r = x & p->half_mask;
	for (int k = 0; k < 4; k++) {
		uint64_t t = l ^ (hj_mix64(r ^ p->round_key[k]) & p->half_mask);
		l = r;
		r = t;
	}
	return (l << p->half_bits) | r;
}

/* bijection [0, n) -> [0, n): cycle-walk the Feistel permutation of the
 * enclosing 2^(2*half_bits) domain (< 4n, so < 4 steps expected) */
HJ_HD uint64_t hj_perm_apply(const hj_perm_t *p, uint64_t x) {
	do {
		x = hj_perm_feistel(p, x);
	} while (x >= p->n);
	return x;
}

/* ---- relation specification ---------------------------------------------- */
typedef struct {
	uint32_t workload;   /* HJ_WL_* */
	uint32_t width;      /* tuple bytes, multiple of 16, >= 16 */
	uint64_t seed;
	uint64_t r_num;      /* |R| */
	uint64_t s_num;      /* |S| */
	uint64_t key_range;  /* HJ_WL_RANDOM: keys in [1, key_range] */
	double theta;        /* HJ_WL_ZIPF exponent */
} hj_gen_spec_t;

typedef struct {
	hj_gen_spec_t spec;
	hj_perm_t r_perm;    /* R key permutation */
	hj_perm_t s_perm;    /* S shuffle (reference-style workloads) */
	hj_perm_t z_perm;    /* Zipf rank -> key */
} hj_gen_state_t;

HJ_HD hj_gen_state_t hj_gen_state(const hj_gen_spec_t *spec) {
	hj_gen_state_t g;
	g.spec = *spec;
	g.r_perm = hj_perm_make(spec->r_num ? spec->r_num : 1, spec->seed, HJ_STREAM_R_PERM);
	g.s_perm = hj_perm_make(spec->s_num ? spec->s_num : 1, spec->seed, HJ_STREAM_S_PERM);
	g.z_perm = hj_perm_make(spec->r_num ? spec->r_num : 1, spec->seed, HJ_STREAM_ZIPF_PERM);
	return g;
}

/* Zipf rank in [0, n) for a uniform double u in [0,1): smallest r with
 * cdf[r] > u, where cdf[r] = sum_{k<=r+1} k^-theta / H (cdf[n-1] forced to 1). */
HJ_HD uint64_t hj_zipf_rank(const double *cdf, uint64_t n, double u) {
	uint64_t lo = 0, hi = n - 1;
	while (lo < hi) {
		uint64_t mid = lo + ((hi - lo) >> 1);
		if (cdf[mid] > u) hi = mid; else lo = mid + 1;
	}
	return lo;
}

HJ_HD double hj_unit_double(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

#include <math.h>
/* Host-only: cumulative Zipf distribution over ranks 1..n with exponent
 * theta, summed in rank order in IEEE double (no reassociation), normalised,
 * last entry forced to 1.0.  The device samples from a copy of this table. */
static inline void hj_zipf_cdf_build(double *cdf, uint64_t n, double theta) {
	double sum = 0.0;
	for (uint64_t k = 1; k <= n; k++) {
		sum += pow((double)k, -theta);
		cdf[k - 1] = sum;
	}
	for (uint64_t k = 0; k < n; k++) cdf[k] = cdf[k] / sum;
	cdf[n - 1] = 1.0;
}

/* Words of R tuple i: out[0] = key, out[1] = payload, out[2..width/8) = rest */
HJ_HD void hj_gen_r_tuple(const hj_gen_state_t *g, uint64_t i, uint64_t *out) {
	const hj_gen_spec_t *s = &g->spec;
	uint64_t key;
	if (s->workload == HJ_WL_RANDOM)
		key = 1u + hj_bounded(hj_rand(s->seed, HJ_STREAM_R_KEY, i), s->key_range);
	else
		key = hj_perm_apply(&g->r_perm, i) + 1u;
	out[0] = key;
	switch (s->workload) {
	case HJ_WL_PAYLOAD_REF: out[1] = i + 1u; break;
	case HJ_WL_ZIPF:
	case HJ_WL_RANDOM: out[1] = hj_rand(s->seed, HJ_STREAM_R_PAY, i); break;
	default: out[1] = key; break;
	}
	for (uint32_t w = 2; w < s->width / 8u; w++) out[w] = hj_rand(s->seed, HJ_STREAM_R_REST, i * 64u + w);
}

/* Words of S tuple j; `zipf_cdf` is only read for HJ_WL_ZIPF. */
HJ_HD void hj_gen_s_tuple(const hj_gen_state_t *g, uint64_t j, const double *zipf_cdf, uint64_t *out) {
	const hj_gen_spec_t *s = &g->spec;
	uint64_t key;
	switch (s->workload) {
	case HJ_WL_PKFK_REF:
	case HJ_WL_PAYLOAD_REF:
		key = (hj_perm_apply(&g->s_perm, j) % s->r_num) + 1u;
		break;
	case HJ_WL_ZIPF: {
		uint64_t rank = hj_zipf_rank(zipf_cdf, s->r_num, hj_unit_double(hj_rand(s->seed, HJ_STREAM_S_KEY, j)));
		key = hj_perm_apply(&g->z_perm, rank) + 1u;
		break;
	}
	case HJ_WL_RANDOM:
		key = 1u + hj_bounded(hj_rand(s->seed, HJ_STREAM_S_KEY, j), s->key_range);
		break;
	default:
		key = 1u + hj_bounded(hj_rand(s->seed, HJ_STREAM_S_KEY, j), s->r_num);
		break;
	}
	out[0] = key;
	switch (s->workload) {
	case HJ_WL_PKFK_REF: out[1] = key; break;
	case HJ_WL_ZIPF:
	case HJ_WL_RANDOM: out[1] = hj_rand(s->seed, HJ_STREAM_S_PAY, j); break;
	default: out[1] = j + 1u; break;
	}
	for (uint32_t w = 2; w < s->width / 8u; w++) out[w] = hj_rand(s->seed, HJ_STREAM_S_REST, j * 64u + w);
}

#endif /* HJGEN_H */
