/*
 * hjoracle.c -- TEST INFRASTRUCTURE ONLY (see hjoracle.h).  CPU restatement
 * of the reference's DRAM-mode equi-joins plus independent exact checkers.
 *
 * Followed files (reference paths):
 *   src/algo/nphj.c   SC: build :132-301, probe :304-442, driver :677-1037
 *                     LP: insert :1056-1070, lookup :1146-1167, probe
 *                     :1176-1269, driver :1487-1848
 *   src/algo/nphj.h   bucket_t :14-26, bucket_buffer_t :35-40, hash_entry_t :43-56
 *   src/algo/phj.c    partition :257-359, SWWCB flush :218-254, pass 2
 *                     :366-491, bc_join :496-571, lp_join :574-680,
 *                     drivers :1041-1872, :1875-2534
 *   src/inc/task_pthread.h  mutex LIFO task queue :110-190
 *   src/inc/lock.h    TAS spinlock :16-47
 *   src/inc/store.h   tuple_t :87-96, SWWCB_SIZE = 64 in DRAM mode :159-165
 *   src/algo/params.h IDHASH :9-11, FANOUT_PASS1/2 :34-35
 */
#define _GNU_SOURCE
#include "hjoracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define BUCKET_CAPACITY 2       /* params.h:16 */
#define OVERFLOW_BUF_SIZE 1024  /* params.h:21 */
#define SWWCB_SIZE 64           /* store.h:159-165, DRAM */

static double now_s(void) {
	struct timespec t;
	clock_gettime(CLOCK_REALTIME, &t);
	return t.tv_sec + t.tv_nsec * 1e-9;
}

static inline uint64_t key_at(const uint8_t *base, uint64_t i, uint32_t w) {
	uint64_t k;
	memcpy(&k, base + i * w, 8);
	return k;
}
static inline uint64_t rid_at(const uint8_t *base, uint64_t i, uint32_t w) {
	uint64_t r;
	memcpy(&r, base + i * w + 8, 8);
	return r;
}
/* integer next power of two (utils.c:351-353 computes it in floating point,
 * which rounds some exact powers up; that only changes table sizes) */
static uint64_t np2(uint64_t x) {
	uint64_t p = 1;
	while (p < x) p <<= 1;
	return p;
}

static void slice(uint64_t n, int t, int nt, uint64_t *a, uint64_t *b) {
	uint64_t per = n / (uint64_t)nt;
	*a = per * (uint64_t)t;
	*b = (t == nt - 1) ? n : *a + per;
}

/* ------------------------------------------------------------ generator */
typedef struct {
	const hj_gen_spec_t *spec;
	hj_gen_state_t g;
	const double *cdf;
	uint8_t *r, *s;
	int t, nt;
} gen_arg_t;

static void *gen_thread(void *p) {
	gen_arg_t *a = (gen_arg_t *)p;
	uint64_t w[64], lo, hi;
	const uint32_t W = a->spec->width;
	slice(a->spec->r_num, a->t, a->nt, &lo, &hi);
	for (uint64_t i = lo; i < hi; i++) {
		hj_gen_r_tuple(&a->g, i, w);
		memcpy(a->r + i * W, w, W);
	}
	slice(a->spec->s_num, a->t, a->nt, &lo, &hi);
	for (uint64_t j = lo; j < hi; j++) {
		hj_gen_s_tuple(&a->g, j, a->cdf, w);
		memcpy(a->s + j * W, w, W);
	}
	return NULL;
}

void or_generate(const hj_gen_spec_t *spec, void *r, void *s, int threads) {
	if (threads < 1) threads = 1;
	double *cdf = NULL;
	if (spec->workload == HJ_WL_ZIPF && spec->s_num) {
		cdf = (double *)malloc(spec->r_num * sizeof(double));
		hj_zipf_cdf_build(cdf, spec->r_num, spec->theta);
	}
	pthread_t th[256];
	gen_arg_t args[256];
	if (threads > 256) threads = 256;
	for (int t = 0; t < threads; t++) {
		args[t] = (gen_arg_t){spec, hj_gen_state(spec), cdf, (uint8_t *)r, (uint8_t *)s, t, threads};
		pthread_create(&th[t], NULL, gen_thread, &args[t]);
	}
	for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
	free(cdf);
}

/* ------------------------------------------------------------ nphj_sc */
typedef struct bucket_s {
	volatile char latch;         /* lock.h TAS latch */
	int count;
	struct bucket_s *next;
	uint8_t tuples[];            /* BUCKET_CAPACITY tuples of W bytes */
} bucket_t;

typedef struct bucket_buffer_s {
	struct bucket_buffer_s *next;
	int count;
	uint8_t *buf;                /* OVERFLOW_BUF_SIZE buckets */
} bucket_buffer_t;

static inline void spin_lock(volatile char *l) {
	while (__sync_lock_test_and_set(l, 1))
		while (*l) __builtin_ia32_pause();
}
static inline void spin_unlock(volatile char *l) { __sync_lock_release(l); }

typedef struct {
	int tid, nt;
	pthread_barrier_t *bar;
	const uint8_t *R, *S;
	uint64_t nr, ns;
	uint32_t W;
	uint8_t *buckets;
	size_t bsize;
	uint64_t mask;
	bucket_buffer_t *ovf;
	uint64_t cnt, sum;
	double t0, t1, t2;
} sc_arg_t;

static bucket_t *new_overflow_bucket(bucket_buffer_t **pb, size_t bsize) {
	bucket_buffer_t *b = *pb;
	if (!b || b->count == OVERFLOW_BUF_SIZE) {
		bucket_buffer_t *nb = (bucket_buffer_t *)malloc(sizeof *nb);
		nb->buf = (uint8_t *)calloc(OVERFLOW_BUF_SIZE, bsize);
		nb->count = 0;
		nb->next = b;
		*pb = b = nb;
	}
	return (bucket_t *)(b->buf + bsize * (size_t)b->count++);
}

static void *sc_thread(void *p) {
	sc_arg_t *a = (sc_arg_t *)p;
	uint64_t lo, hi;
	const uint32_t W = a->W;
	pthread_barrier_wait(a->bar);
	if (a->tid == 0) a->t0 = now_s();
	/* build: nphj.c:235-267 */
	slice(a->nr, a->tid, a->nt, &lo, &hi);
	for (uint64_t i = lo; i < hi; i++) {
		const uint8_t *t = a->R + i * W;
		bucket_t *curr = (bucket_t *)(a->buckets + a->bsize * (key_at(a->R, i, W) & a->mask));
		uint8_t *dest;
		spin_lock(&curr->latch);
		bucket_t *nxt = curr->next;
		if (curr->count == BUCKET_CAPACITY) {
			if (!nxt || nxt->count == BUCKET_CAPACITY) {
				bucket_t *b = new_overflow_bucket(&a->ovf, a->bsize);
				curr->next = b;
				b->next = nxt;
				b->count = 1;
				dest = b->tuples;
			} else {
				dest = nxt->tuples + (size_t)W * nxt->count;
				nxt->count++;
			}
		} else {
			dest = curr->tuples + (size_t)W * curr->count;
			curr->count++;
		}
		spin_unlock(&curr->latch);
		memcpy(dest, t, W);
	}
	pthread_barrier_wait(a->bar);
	if (a->tid == 0) a->t1 = now_s();
	/* probe: nphj.c:393-417 */
	slice(a->ns, a->tid, a->nt, &lo, &hi);
	uint64_t cnt = 0, sum = 0;
	for (uint64_t j = lo; j < hi; j++) {
		const uint64_t sk = key_at(a->S, j, W), sr = rid_at(a->S, j, W);
		const bucket_t *b = (const bucket_t *)(a->buckets + a->bsize * (sk & a->mask));
		do {
			for (int k = 0; k < b->count; k++) {
				const uint64_t f = key_at(b->tuples, (uint64_t)k, W) == sk;
				cnt += f;
				sum += (sr + rid_at(b->tuples, (uint64_t)k, W)) * f;
			}
			b = b->next;
		} while (b);
	}
	a->cnt = cnt;
	a->sum = sum;
	pthread_barrier_wait(a->bar);
	if (a->tid == 0) a->t2 = now_s();
	return NULL;
}

int or_nphj_sc(const void *R, uint64_t nr, const void *S, uint64_t ns, uint32_t width, int threads,
               or_result_t *out) {
	if (threads < 1) threads = 1;
	if (threads > 256) threads = 256;
	uint64_t nb = nr / BUCKET_CAPACITY; /* nphj.c:746 */
	if (nb == 0) nb = 1;
	const size_t bsize = (sizeof(bucket_t) + (size_t)BUCKET_CAPACITY * width + 63) & ~(size_t)63;
	uint8_t *buckets = (uint8_t *)aligned_alloc(64, bsize * nb);
	if (!buckets) return -2;
	memset(buckets, 0, bsize * nb); /* memset before the timed build, nphj.c:452-460 */
	pthread_barrier_t bar;
	pthread_barrier_init(&bar, NULL, (unsigned)threads);
	pthread_t th[256];
	sc_arg_t *args = (sc_arg_t *)calloc((size_t)threads, sizeof *args);
	for (int t = 0; t < threads; t++) {
		args[t] = (sc_arg_t){t, threads, &bar, (const uint8_t *)R, (const uint8_t *)S, nr, ns, width,
		                     buckets, bsize, nb - 1, NULL, 0, 0, 0, 0, 0};
		pthread_create(&th[t], NULL, sc_thread, &args[t]);
	}
	uint64_t cnt = 0, sum = 0;
	for (int t = 0; t < threads; t++) {
		pthread_join(th[t], NULL);
		cnt += args[t].cnt;
		sum += args[t].sum;
	}
	out->matched_cnt = cnt;
	out->check_sum = sum;
	out->buildpart_time = args[0].t1 - args[0].t0;
	out->probejoin_time = args[0].t2 - args[0].t1;
	out->total_time = args[0].t2 - args[0].t0;
	for (int t = 0; t < threads; t++) {
		bucket_buffer_t *b = args[t].ovf;
		while (b) {
			bucket_buffer_t *n = b->next;
			free(b->buf);
			free(b);
			b = n;
		}
	}
	free(args);
	free(buckets);
	pthread_barrier_destroy(&bar);
	return 0;
}

/* ------------------------------------------------------------ nphj_lp */
typedef struct {
	int tid, nt;
	pthread_barrier_t *bar;
	const uint8_t *R, *S;
	uint64_t nr, ns;
	uint32_t W;
	uint8_t *table;  /* entries: u64 hashkey + W-byte tuple */
	size_t esize;
	uint64_t mask;
	uint64_t cnt, sum;
	double t0, t1, t2;
} lp_arg_t;

static void *lp_thread(void *p) {
	lp_arg_t *a = (lp_arg_t *)p;
	uint64_t lo, hi;
	const uint32_t W = a->W;
	pthread_barrier_wait(a->bar);
	if (a->tid == 0) a->t0 = now_s();
	/* build: lp_insert_with_pos, nphj.c:1056-1070 */
	slice(a->nr, a->tid, a->nt, &lo, &hi);
	for (uint64_t i = lo; i < hi; i++) {
		const uint64_t h = key_at(a->R, i, W) & a->mask;
		uint64_t pos = h;
		for (;;) {
			uint64_t *hk = (uint64_t *)(a->table + a->esize * pos);
			if (*hk == 0 && __sync_bool_compare_and_swap(hk, 0, h + 1)) break;
			pos = (pos + 1) & a->mask;
		}
		memcpy(a->table + a->esize * pos + 8, a->R + i * W, W);
	}
	pthread_barrier_wait(a->bar);
	if (a->tid == 0) a->t1 = now_s();
	/* probe: lp_lookup_with_pos, nphj.c:1162-1165, bounded as :1147-1160 */
	slice(a->ns, a->tid, a->nt, &lo, &hi);
	uint64_t cnt = 0, sum = 0;
	for (uint64_t j = lo; j < hi; j++) {
		const uint64_t sk = key_at(a->S, j, W), sr = rid_at(a->S, j, W);
		uint64_t pos = sk & a->mask;
		for (uint64_t step = 0; step <= a->mask; step++) {
			const uint8_t *e = a->table + a->esize * pos;
			if (*(const uint64_t *)e == 0) break;
			if (key_at(e + 8, 0, W) == sk) {
				cnt++;
				sum += sr + rid_at(e + 8, 0, W);
				break;
			}
			pos = (pos + 1) & a->mask;
		}
	}
	a->cnt = cnt;
	a->sum = sum;
	pthread_barrier_wait(a->bar);
	if (a->tid == 0) a->t2 = now_s();
	return NULL;
}

int or_nphj_lp(const void *R, uint64_t nr, const void *S, uint64_t ns, uint32_t width, int threads,
               or_result_t *out) {
	if (threads < 1) threads = 1;
	if (threads > 256) threads = 256;
	const uint64_t slots = np2(nr ? nr : 1); /* nphj.c:1556 */
	const size_t esize = (8 + width + 15) & ~(size_t)15;
	uint8_t *table = (uint8_t *)aligned_alloc(64, (esize * slots + 63) & ~(size_t)63);
	if (!table) return -2;
	memset(table, 0, esize * slots);
	pthread_barrier_t bar;
	pthread_barrier_init(&bar, NULL, (unsigned)threads);
	pthread_t th[256];
	lp_arg_t *args = (lp_arg_t *)calloc((size_t)threads, sizeof *args);
	for (int t = 0; t < threads; t++) {
		args[t] = (lp_arg_t){t, threads, &bar, (const uint8_t *)R, (const uint8_t *)S, nr, ns, width,
		                     table, esize, slots - 1, 0, 0, 0, 0, 0};
		pthread_create(&th[t], NULL, lp_thread, &args[t]);
	}
	uint64_t cnt = 0, sum = 0;
	for (int t = 0; t < threads; t++) {
		pthread_join(th[t], NULL);
		cnt += args[t].cnt;
		sum += args[t].sum;
	}
	out->matched_cnt = cnt;
	out->check_sum = sum;
	out->buildpart_time = args[0].t1 - args[0].t0;
	out->probejoin_time = args[0].t2 - args[0].t1;
	out->total_time = args[0].t2 - args[0].t0;
	free(args);
	free(table);
	pthread_barrier_destroy(&bar);
	return 0;
}

/* ------------------------------------------------------------ radix */
/* One pass of the parallel radix partition with SWWCBs (phj.c:257-359):
 * thread t owns input slice t; hist[t][p] = count, prefix across threads in
 * thread order (stable); scatter through 64-byte line buffers whose full
 * lines are stored at line-aligned output positions; after a barrier every
 * thread writes the remainder of its buffers. */
typedef struct {
	int nt;
	pthread_barrier_t *bar;
	const uint8_t *in;
	uint64_t n;
	uint8_t *out;
	uint32_t W, shift, bits;
	uint64_t **hist; /* [thread][F] */
} part_shared_t;

static void partition_slice(part_shared_t *sh, int tid) {
	const uint32_t W = sh->W, F = 1u << sh->bits;
	const uint64_t fmask = F - 1;
	uint64_t lo, hi;
	slice(sh->n, tid, sh->nt, &lo, &hi);
	uint64_t *myhist = sh->hist[tid];
	memset(myhist, 0, sizeof(uint64_t) * F);
	for (uint64_t i = lo; i < hi; i++) myhist[(key_at(sh->in, i, W) >> sh->shift) & fmask]++;
	pthread_barrier_wait(sh->bar);
	/* output_offset[p] = sum over partitions < p of all threads + sum of
	 * threads < tid for p  (phj.c:299-311) */
	uint64_t *off = (uint64_t *)malloc(sizeof(uint64_t) * F);
	uint64_t run = 0;
	for (uint32_t p = 0; p < F; p++) {
		uint64_t before = 0, all = 0;
		for (int t = 0; t < sh->nt; t++) {
			if (t < tid) before += sh->hist[t][p];
			all += sh->hist[t][p];
		}
		off[p] = run + before;
		run += all;
	}
	const uint32_t tps = W <= SWWCB_SIZE ? SWWCB_SIZE / W : 1; /* TUPLE_PER_SWWCB */
	uint8_t *buf = (uint8_t *)aligned_alloc(64, (size_t)F * tps * W);
	uint64_t *slot = (uint64_t *)malloc(sizeof(uint64_t) * F);
	for (uint32_t p = 0; p < F; p++) slot[p] = off[p];
	for (uint64_t i = lo; i < hi; i++) {
		const uint32_t p = (uint32_t)((key_at(sh->in, i, W) >> sh->shift) & fmask);
		const uint64_t s = slot[p], m = s % tps;
		memcpy(buf + ((size_t)p * tps + m) * W, sh->in + i * W, W);
		if (m == tps - 1) {
			/* full line: store the whole buffer at its aligned position; a
			 * leading part before off[p] belongs to the previous slice and is
			 * rewritten by that slice's remainder pass */
			const uint64_t first = s - (tps - 1);
			for (uint32_t k = 0; k < tps; k++)
				if (first + k >= off[p]) memcpy(sh->out + (first + k) * W, buf + ((size_t)p * tps + k) * W, W);
		}
		slot[p] = s + 1;
	}
	pthread_barrier_wait(sh->bar);
	for (uint32_t p = 0; p < F; p++) {
		const uint64_t s = slot[p], m = s % tps, first = s - m;
		for (uint64_t k = 0; k < m; k++)
			if (first + k >= off[p]) memcpy(sh->out + (first + k) * W, buf + ((size_t)p * tps + k) * W, W);
	}
	free(buf);
	free(slot);
	free(off);
}

typedef struct {
	part_shared_t *sh;
	int tid;
} part_arg_t;

static void *part_thread(void *p) {
	part_arg_t *a = (part_arg_t *)p;
	partition_slice(a->sh, a->tid);
	return NULL;
}

int or_radix_partition(const void *in, uint64_t n, uint32_t width, uint32_t shift, uint32_t bits, int threads,
                       void *out, uint64_t *hist) {
	if (threads < 1) threads = 1;
	if (threads > 256) threads = 256;
	const uint32_t F = 1u << bits;
	pthread_barrier_t bar;
	pthread_barrier_init(&bar, NULL, (unsigned)threads);
	uint64_t **h = (uint64_t **)malloc(sizeof(uint64_t *) * threads);
	for (int t = 0; t < threads; t++) h[t] = (uint64_t *)malloc(sizeof(uint64_t) * F);
	part_shared_t sh = {threads, &bar, (const uint8_t *)in, n, (uint8_t *)out, width, shift, bits, h};
	pthread_t th[256];
	part_arg_t args[256];
	for (int t = 0; t < threads; t++) {
		args[t] = (part_arg_t){&sh, t};
		pthread_create(&th[t], NULL, part_thread, &args[t]);
	}
	for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
	if (hist)
		for (uint32_t p = 0; p < F; p++) {
			hist[p] = 0;
			for (int t = 0; t < threads; t++) hist[p] += h[t][p];
		}
	for (int t = 0; t < threads; t++) free(h[t]);
	free(h);
	pthread_barrier_destroy(&bar);
	return 0;
}

/* serial sub-partition of one pass-1 partition (phj.c:366-491): stable
 * counting sort by bits [shift, shift+bits) into `out` */
static void serial_partition(const uint8_t *in, uint64_t n, uint32_t W, uint32_t shift, uint32_t bits,
                             uint8_t *out, uint64_t *hist /* F+1 */) {
	const uint32_t F = 1u << bits;
	const uint64_t fmask = F - 1;
	memset(hist, 0, sizeof(uint64_t) * (F + 1));
	for (uint64_t i = 0; i < n; i++) hist[((key_at(in, i, W) >> shift) & fmask) + 1]++;
	for (uint32_t p = 0; p < F; p++) hist[p + 1] += hist[p];
	uint64_t *cur = (uint64_t *)malloc(sizeof(uint64_t) * F);
	memcpy(cur, hist, sizeof(uint64_t) * F);
	for (uint64_t i = 0; i < n; i++) {
		const uint32_t p = (uint32_t)((key_at(in, i, W) >> shift) & fmask);
		memcpy(out + cur[p]++ * W, in + i * W, W);
	}
	free(cur);
}

/* task queue (task_pthread.h:110-133): mutex-protected LIFO */
typedef struct {
	uint64_t r0, r1, s0, s1;  /* ranges into the current partitioned buffers */
} task_t;

typedef struct {
	pthread_mutex_t mu;
	task_t *tasks;
	uint64_t n, cap;
} task_queue_t;

static void tq_init(task_queue_t *q, uint64_t cap) {
	pthread_mutex_init(&q->mu, NULL);
	q->tasks = (task_t *)malloc(sizeof(task_t) * (cap ? cap : 1));
	q->n = 0;
	q->cap = cap ? cap : 1;
}
static void tq_add(task_queue_t *q, task_t t) {
	pthread_mutex_lock(&q->mu);
	if (q->n == q->cap) {
		q->cap *= 2;
		q->tasks = (task_t *)realloc(q->tasks, sizeof(task_t) * q->cap);
	}
	q->tasks[q->n++] = t;
	pthread_mutex_unlock(&q->mu);
}
static int tq_get(task_queue_t *q, task_t *t) {
	int ok = 0;
	pthread_mutex_lock(&q->mu);
	if (q->n) {
		*t = q->tasks[--q->n];
		ok = 1;
	}
	pthread_mutex_unlock(&q->mu);
	return ok;
}

typedef struct {
	int nt, lp;  /* lp: per-partition join kind (0 bc, 1 lp, 2 sc, 3 hr, 4 hro) */
	pthread_barrier_t *bar;
	part_shared_t pr, ps;        /* pass-1 partition of R and of S */
	uint32_t W, B, b1, b2, passes;
	uint8_t *r1, *s1, *r2, *s2;  /* pass-1 / pass-2 outputs */
	uint64_t nr, ns;
	task_queue_t part_q, join_q;
	double t0, t1, t2;
} radix_shared_t;

typedef struct {
	radix_shared_t *sh;
	int tid;
	uint64_t cnt, sum;
} radix_arg_t;

/* bc_join, phj.c:496-571: heads u64[np2(|R_p|)], next u64[|R_p|], 1-based,
 * head insertion; hash (key >> B) & (np2 - 1) */
static void bc_join(const uint8_t *R, uint64_t nr, const uint8_t *S, uint64_t ns, uint32_t W, uint32_t B,
                    uint64_t *scratch, uint64_t *cnt, uint64_t *sum) {
	const uint64_t nb = np2(nr), mask = nb - 1;
	uint64_t *bucket = scratch, *next = scratch + nb;
	memset(scratch, 0, sizeof(uint64_t) * (nb + nr));
	for (uint64_t i = 0; i < nr;) {
		const uint64_t h = (key_at(R, i, W) >> B) & mask;
		next[i] = bucket[h];
		bucket[h] = ++i;
	}
	uint64_t c = 0, s = 0;
	for (uint64_t j = 0; j < ns; j++) {
		const uint64_t sk = key_at(S, j, W), sr = rid_at(S, j, W);
		for (uint64_t hit = bucket[(sk >> B) & mask]; hit > 0; hit = next[hit - 1]) {
			const uint64_t f = key_at(R, hit - 1, W) == sk;
			c += f;
			s += (rid_at(R, hit - 1, W) + sr) * f;
		}
	}
	*cnt += c;
	*sum += s;
}

/* lp_join, phj.c:574-680: np2(|R_p|) slots filled in R_p order; occupancy
 * kept in a separate array (the reference tests key != 0) */
static void lp_join(const uint8_t *R, uint64_t nr, const uint8_t *S, uint64_t ns, uint32_t W, uint32_t B,
                    uint64_t *scratch, uint64_t *cnt, uint64_t *sum) {
	const uint64_t nb = np2(nr), mask = nb - 1;
	uint64_t *occ = scratch; /* R index + 1 per slot */
	memset(occ, 0, sizeof(uint64_t) * nb);
	for (uint64_t i = 0; i < nr; i++) {
		uint64_t pos = (key_at(R, i, W) >> B) & mask;
		while (occ[pos]) pos = (pos + 1) & mask;
		occ[pos] = i + 1;
	}
	uint64_t c = 0, s = 0;
	for (uint64_t j = 0; j < ns; j++) {
		const uint64_t sk = key_at(S, j, W), sr = rid_at(S, j, W);
		uint64_t pos = (sk >> B) & mask;
		for (uint64_t step = 0; step < nb; step++) {
			const uint64_t o = occ[pos];
			if (!o) break;
			if (key_at(R, o - 1, W) == sk) {
				c++;
				s += sr + rid_at(R, o - 1, W);
				break;
			}
			pos = (pos + 1) & mask;
		}
	}
	*cnt += c;
	*sum += s;
}

/* sc_join, phj.c:683-781: np2(|R_p|)/2 buckets of PHJ_SC_BUCKET_CAPACITY = 2
 * tuples (CMakeLists.txt:330-333) with a count and a next pointer; a full
 * bucket gets an overflow bucket inserted right after it (or the existing
 * one is filled); the probe compares every tuple of the chain.  Departure: at
 * least one bucket (for |R_p| = 1 the reference's mask underflows). */
typedef struct or_sc_bkt {
	uint32_t count;
	struct or_sc_bkt *next;
	uint64_t key[2], rid[2];
} or_sc_bkt_t;

static void sc_join(const uint8_t *R, uint64_t nr, const uint8_t *S, uint64_t ns, uint32_t W, uint32_t B,
                    uint64_t *cnt, uint64_t *sum) {
	uint64_t nbk = np2(nr) / 2;
	if (nbk == 0) nbk = 1;
	const uint64_t mask = nbk - 1;
	or_sc_bkt_t *bk = (or_sc_bkt_t *)calloc(nbk * 2, sizeof(or_sc_bkt_t));
	or_sc_bkt_t *pool = bk + nbk;
	uint64_t free_idx = 0;
	for (uint64_t i = 0; i < nr; i++) {
		or_sc_bkt_t *curr = bk + ((key_at(R, i, W) >> B) & mask), *nxt = curr->next, *dest;
		if (curr->count == 2) {
			if (!nxt || nxt->count == 2) {
				dest = pool + free_idx++;
				curr->next = dest;
				dest->next = nxt;
			} else {
				dest = nxt;
			}
		} else {
			dest = curr;
		}
		dest->key[dest->count] = key_at(R, i, W);
		dest->rid[dest->count] = rid_at(R, i, W);
		dest->count++;
	}
	uint64_t c = 0, s = 0;
	for (uint64_t j = 0; j < ns; j++) {
		const uint64_t sk = key_at(S, j, W), sr = rid_at(S, j, W);
		for (const or_sc_bkt_t *t = bk + ((sk >> B) & mask); t; t = t->next)
			for (uint32_t k = 0; k < t->count; k++) {
				const uint64_t f = t->key[k] == sk;
				c += f;
				s += (t->rid[k] + sr) * f;
			}
	}
	free(bk);
	*cnt += c;
	*sum += s;
}

/* hr_join (phj.c:784-877, shrink = 0) and hro_join (phj.c:881-1037, shrink =
 * log2(64 / W), at least 2^shrink buckets): count R_p per bucket, exclusive
 * prefix, scatter R_p into bucket order (the running counters end as the
 * bucket ends), probe the bucket's range.  hro reads the range in blocks of
 * 64 / W tuples as its 512-bit compare does; tuples past the range belong to
 * other buckets and never match.  Departure: the zero padding past |R_p| is
 * not compared (in the reference a probe key 0 can match it). */
static void hr_join(const uint8_t *R, uint64_t nr, const uint8_t *S, uint64_t ns, uint32_t W, uint32_t B,
                    int shrink, uint64_t *cnt, uint64_t *sum) {
	uint64_t nb = np2(nr) >> shrink;
	if (nb < (1ull << shrink)) nb = 1ull << shrink;
	const uint64_t mask = nb - 1;
	uint64_t *hist = (uint64_t *)calloc(nb, 8);
	uint8_t *tmp = (uint8_t *)malloc(nr * W + 1);
	for (uint64_t i = 0; i < nr; i++) hist[(key_at(R, i, W) >> B) & mask]++;
	uint64_t acc = 0;
	for (uint64_t h = 0; h < nb; h++) {
		acc += hist[h];
		hist[h] = acc - hist[h];
	}
	for (uint64_t i = 0; i < nr; i++) {
		const uint64_t h = (key_at(R, i, W) >> B) & mask;
		memcpy(tmp + hist[h] * W, R + i * W, W);
		hist[h]++;
	}
	const uint64_t blk = shrink ? 64 / W : 1;
	uint64_t c = 0, s = 0;
	for (uint64_t j = 0; j < ns; j++) {
		const uint64_t sk = key_at(S, j, W), sr = rid_at(S, j, W), h = (sk >> B) & mask;
		const uint64_t st = h ? hist[h - 1] : 0, en = hist[h];
		for (uint64_t t0 = st; t0 < en; t0 += blk)
			for (uint64_t t = t0; t < t0 + blk && t < nr; t++) {
				const uint64_t f = key_at(tmp, t, W) == sk;
				c += f;
				s += (rid_at(tmp, t, W) + sr) * f;
			}
	}
	free(hist);
	free(tmp);
	*cnt += c;
	*sum += s;
}

static void *radix_thread(void *p) {
	radix_arg_t *a = (radix_arg_t *)p;
	radix_shared_t *sh = a->sh;
	const uint32_t W = sh->W;
	pthread_barrier_wait(sh->bar);
	if (a->tid == 0) sh->t0 = now_s();
	partition_slice(&sh->pr, a->tid);  /* pass 1 over R (prj_thread_buildpart) */
	partition_slice(&sh->ps, a->tid);  /* pass 1 over S */
	pthread_barrier_wait(sh->bar);
	if (a->tid == 0) {
		/* thread 0 turns partition pairs into tasks (phj.c:1257-1311) */
		const uint32_t F1 = 1u << sh->b1;
		uint64_t *hr = (uint64_t *)calloc(F1 + 1, 8), *hs = (uint64_t *)calloc(F1 + 1, 8);
		for (uint32_t q = 0; q < F1; q++)
			for (int t = 0; t < sh->nt; t++) {
				hr[q + 1] += sh->pr.hist[t][q];
				hs[q + 1] += sh->ps.hist[t][q];
			}
		for (uint32_t q = 0; q < F1; q++) {
			hr[q + 1] += hr[q];
			hs[q + 1] += hs[q];
		}
		for (uint32_t q = 0; q < F1; q++) {
			task_t t = {hr[q], hr[q + 1], hs[q], hs[q + 1]};
			if (t.r1 > t.r0 && t.s1 > t.s0) tq_add(sh->passes == 2 ? &sh->part_q : &sh->join_q, t);
		}
		free(hr);
		free(hs);
	}
	pthread_barrier_wait(sh->bar);
	if (sh->passes == 2) {
		/* pass 2: each task is sub-partitioned by one thread (phj.c:1344-1349) */
		const uint32_t F2 = 1u << sh->b2;
		uint64_t *hr = (uint64_t *)malloc(sizeof(uint64_t) * (F2 + 1));
		uint64_t *hs = (uint64_t *)malloc(sizeof(uint64_t) * (F2 + 1));
		task_t t;
		while (tq_get(&sh->part_q, &t)) {
			serial_partition(sh->r1 + t.r0 * W, t.r1 - t.r0, W, sh->b1, sh->b2, sh->r2 + t.r0 * W, hr);
			serial_partition(sh->s1 + t.s0 * W, t.s1 - t.s0, W, sh->b1, sh->b2, sh->s2 + t.s0 * W, hs);
			for (uint32_t q = 0; q < F2; q++) {
				task_t u = {t.r0 + hr[q], t.r0 + hr[q + 1], t.s0 + hs[q], t.s0 + hs[q + 1]};
				if (u.r1 > u.r0 && u.s1 > u.s0) tq_add(&sh->join_q, u);
			}
		}
		free(hr);
		free(hs);
	}
	pthread_barrier_wait(sh->bar);
	if (a->tid == 0) sh->t1 = now_s();
	/* join phase (prj_thread_probejoin, phj.c:1746-1872) */
	const uint8_t *RB = sh->passes == 2 ? sh->r2 : sh->r1;
	const uint8_t *SB = sh->passes == 2 ? sh->s2 : sh->s1;
	uint64_t *scratch = NULL;
	size_t scap = 0;
	task_t t;
	while (tq_get(&sh->join_q, &t)) {
		const uint64_t nr = t.r1 - t.r0;
		const size_t need = (np2(nr) + nr + 1) * sizeof(uint64_t);
		if (need > scap) {
			free(scratch);
			scratch = (uint64_t *)malloc(need);
			scap = need;
		}
		const uint8_t *RP = RB + t.r0 * W, *SP = SB + t.s0 * W;
		const uint64_t ns = t.s1 - t.s0;
		switch (sh->lp) {  /* join kind: 0 bc, 1 lp, 2 sc, 3 hr, 4 hro */
		case 1: lp_join(RP, nr, SP, ns, W, sh->B, scratch, &a->cnt, &a->sum); break;
		case 2: sc_join(RP, nr, SP, ns, W, sh->B, &a->cnt, &a->sum); break;
		case 3: hr_join(RP, nr, SP, ns, W, sh->B, 0, &a->cnt, &a->sum); break;
		case 4: hr_join(RP, nr, SP, ns, W, sh->B, W == 16 ? 2 : (W == 32 ? 1 : 0), &a->cnt, &a->sum); break;
		default: bc_join(RP, nr, SP, ns, W, sh->B, scratch, &a->cnt, &a->sum); break;
		}
	}
	free(scratch);
	pthread_barrier_wait(sh->bar);
	if (a->tid == 0) sh->t2 = now_s();
	return NULL;
}

static int radix_join(const void *R, uint64_t nr, const void *S, uint64_t ns, uint32_t width, uint32_t radix_bits,
                      uint32_t passes, uint32_t b1, int kind, int threads, or_result_t *out);

int or_phj_radix(const void *R, uint64_t nr, const void *S, uint64_t ns, uint32_t width, uint32_t radix_bits,
                 uint32_t passes, int kind, int threads, or_result_t *out) {
	if (passes < 1 || passes > 2) return -1;
	return radix_join(R, nr, S, ns, width, radix_bits, passes, radix_bits / passes, kind, threads, out);
}

/* phj_asymm_radix_* (src/algo/phj_asymm_radix.c): the first pass takes
 * NUM_RADIX_BITS_PASS1 bits, the second the rest (phj_asymm_radix.h:10-22);
 * a first pass of all the bits leaves a single pass */
int or_phj_asymm_radix(const void *R, uint64_t nr, const void *S, uint64_t ns, uint32_t width, uint32_t radix_bits,
                       uint32_t radix_bits_pass1, int kind, int threads, or_result_t *out) {
	if (radix_bits_pass1 >= radix_bits)
		return radix_join(R, nr, S, ns, width, radix_bits, 1, radix_bits, kind, threads, out);
	return radix_join(R, nr, S, ns, width, radix_bits, 2, radix_bits_pass1, kind, threads, out);
}

static int radix_join(const void *R, uint64_t nr, const void *S, uint64_t ns, uint32_t width, uint32_t radix_bits,
                      uint32_t passes, uint32_t b1, int kind, int threads, or_result_t *out) {
	if (threads < 1) threads = 1;
	if (threads > 256) threads = 256;
	if (passes < 1 || passes > 2 || radix_bits > 20 || (passes == 2 && radix_bits < 2)) return -1;
	if (passes == 2 && (b1 < 1 || b1 >= radix_bits)) return -1;
	radix_shared_t sh;
	memset(&sh, 0, sizeof sh);
	sh.nt = threads;
	if (kind < 0 || kind > 4) return -1;
	sh.lp = kind;
	sh.W = width;
	sh.B = radix_bits;
	sh.passes = passes;
	sh.b1 = passes == 2 ? b1 : radix_bits;
	sh.b2 = radix_bits - sh.b1;
	sh.nr = nr;
	sh.ns = ns;
	pthread_barrier_t bar;
	pthread_barrier_init(&bar, NULL, (unsigned)threads);
	sh.bar = &bar;
	sh.r1 = (uint8_t *)aligned_alloc(64, ((nr ? nr : 1) * width + 63) & ~(size_t)63);
	sh.s1 = (uint8_t *)aligned_alloc(64, ((ns ? ns : 1) * width + 63) & ~(size_t)63);
	if (passes == 2) {
		sh.r2 = (uint8_t *)aligned_alloc(64, ((nr ? nr : 1) * width + 63) & ~(size_t)63);
		sh.s2 = (uint8_t *)aligned_alloc(64, ((ns ? ns : 1) * width + 63) & ~(size_t)63);
	}
	if (!sh.r1 || !sh.s1 || (passes == 2 && (!sh.r2 || !sh.s2))) return -2;
	const uint32_t F1 = 1u << sh.b1;
	uint64_t **hr = (uint64_t **)malloc(sizeof(uint64_t *) * threads);
	uint64_t **hs = (uint64_t **)malloc(sizeof(uint64_t *) * threads);
	for (int t = 0; t < threads; t++) {
		hr[t] = (uint64_t *)malloc(sizeof(uint64_t) * F1);
		hs[t] = (uint64_t *)malloc(sizeof(uint64_t) * F1);
	}
	sh.pr = (part_shared_t){threads, &bar, (const uint8_t *)R, nr, sh.r1, width, 0, sh.b1, hr};
	sh.ps = (part_shared_t){threads, &bar, (const uint8_t *)S, ns, sh.s1, width, 0, sh.b1, hs};
	tq_init(&sh.part_q, F1);
	tq_init(&sh.join_q, 1ull << radix_bits);
	pthread_t th[256];
	radix_arg_t *args = (radix_arg_t *)calloc((size_t)threads, sizeof *args);
	for (int t = 0; t < threads; t++) {
		args[t].sh = &sh;
		args[t].tid = t;
		pthread_create(&th[t], NULL, radix_thread, &args[t]);
	}
	uint64_t cnt = 0, sum = 0;
	for (int t = 0; t < threads; t++) {
		pthread_join(th[t], NULL);
		cnt += args[t].cnt;
		sum += args[t].sum;
	}
	out->matched_cnt = cnt;
	out->check_sum = sum;
	out->buildpart_time = sh.t1 - sh.t0;
	out->probejoin_time = sh.t2 - sh.t1;
	out->total_time = sh.t2 - sh.t0;
	for (int t = 0; t < threads; t++) {
		free(hr[t]);
		free(hs[t]);
	}
	free(hr);
	free(hs);
	free(args);
	free(sh.part_q.tasks);
	free(sh.join_q.tasks);
	free(sh.r1);
	free(sh.s1);
	free(sh.r2);
	free(sh.s2);
	pthread_barrier_destroy(&bar);
	return 0;
}

/* ------------------------------------------------------------ checkers */
typedef struct {
	uint64_t key, idx;
} kidx_t;

static int cmp_kidx(const void *a, const void *b) {
	const kidx_t *x = (const kidx_t *)a, *y = (const kidx_t *)b;
	if (x->key != y->key) return x->key < y->key ? -1 : 1;
	return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

/* Per-key aggregation: count = sum_k cR(k)*cS(k), checksum = sum_k
 * [cS(k)*sumR(k) + cR(k)*sumS(k)]; lp uses cR = 1 and the first R row. */
int or_aggregate(const void *Rv, uint64_t nr, const void *Sv, uint64_t ns, uint32_t W, int lp, or_result_t *out) {
	const uint8_t *R = (const uint8_t *)Rv, *S = (const uint8_t *)Sv;
	kidx_t *a = (kidx_t *)malloc(sizeof(kidx_t) * (nr ? nr : 1));
	if (!a) return -2;
	for (uint64_t i = 0; i < nr; i++) a[i] = (kidx_t){key_at(R, i, W), i};
	qsort(a, nr, sizeof(kidx_t), cmp_kidx);
	/* distinct keys with (count, payload sum, first payload) */
	uint64_t nd = 0;
	uint64_t *dk = (uint64_t *)malloc(8 * (nr ? nr : 1)), *dc = (uint64_t *)malloc(8 * (nr ? nr : 1));
	uint64_t *ds = (uint64_t *)malloc(8 * (nr ? nr : 1)), *df = (uint64_t *)malloc(8 * (nr ? nr : 1));
	for (uint64_t i = 0; i < nr; i++) {
		const uint64_t rp = rid_at(R, a[i].idx, W);
		if (nd && dk[nd - 1] == a[i].key) {
			dc[nd - 1]++;
			ds[nd - 1] += rp;
		} else {
			dk[nd] = a[i].key;
			dc[nd] = 1;
			ds[nd] = rp;
			df[nd] = rp; /* sorted by (key, idx): first row of the key */
			nd++;
		}
	}
	uint64_t cnt = 0, sum = 0;
	for (uint64_t j = 0; j < ns; j++) {
		const uint64_t sk = key_at(S, j, W), sr = rid_at(S, j, W);
		uint64_t lo = 0, hi = nd;
		while (lo < hi) {
			uint64_t mid = (lo + hi) / 2;
			if (dk[mid] < sk) lo = mid + 1; else hi = mid;
		}
		if (lo < nd && dk[lo] == sk) {
			if (lp) {
				cnt += 1;
				sum += df[lo] + sr;
			} else {
				cnt += dc[lo];
				sum += ds[lo] + dc[lo] * sr;
			}
		}
	}
	out->matched_cnt = cnt;
	out->check_sum = sum;
	out->buildpart_time = out->probejoin_time = out->total_time = 0;
	free(a);
	free(dk);
	free(dc);
	free(ds);
	free(df);
	return 0;
}

int or_bruteforce(const void *Rv, uint64_t nr, const void *Sv, uint64_t ns, uint32_t W, int lp, or_result_t *out) {
	const uint8_t *R = (const uint8_t *)Rv, *S = (const uint8_t *)Sv;
	uint64_t cnt = 0, sum = 0;
	for (uint64_t j = 0; j < ns; j++) {
		const uint64_t sk = key_at(S, j, W), sr = rid_at(S, j, W);
		for (uint64_t i = 0; i < nr; i++) {
			if (key_at(R, i, W) == sk) {
				cnt++;
				sum += rid_at(R, i, W) + sr;
				if (lp) break;
			}
		}
	}
	out->matched_cnt = cnt;
	out->check_sum = sum;
	out->buildpart_time = out->probejoin_time = out->total_time = 0;
	return 0;
}

/* ---- late materialisation (src/algo/ptr_nphj.c, ptr_phj.c) ---------------
 * The reference joins {key, row_id} heads, writes one row_id_pair_t
 * {r.row_id, s.row_id} per match, then ptr_retrieve (src/inc/utils.c:
 * 712-723) adds R[r_row_id - 1].row_id + S[s_row_id - 1].row_id over the
 * pairs.  A row id outside 1..n is an error here (the reference reads out of
 * bounds). */

int or_ptr_retrieve(const uint64_t *pairs, uint64_t npairs, const void *Rv, uint64_t nr, const void *Sv,
                    uint64_t ns, uint32_t W, uint64_t *checksum) {
	const uint8_t *R = (const uint8_t *)Rv, *S = (const uint8_t *)Sv;
	uint64_t sum = 0;
	for (uint64_t k = 0; k < npairs; k++) {
		const uint64_t r = pairs[2 * k], s = pairs[2 * k + 1];
		if (r - 1 >= nr || s - 1 >= ns) return -1;
		sum += rid_at(R, r - 1, W) + rid_at(S, s - 1, W);
	}
	*checksum = sum;
	return 0;
}

/* Literal pipeline on small inputs: nested loop emitting the pairs (lp: the
 * first R row of the key), then the retrieve. */
int or_ptr_bruteforce(const void *Rv, uint64_t nr, const void *Sv, uint64_t ns, uint32_t W, int lp,
                      or_result_t *out) {
	const uint8_t *R = (const uint8_t *)Rv, *S = (const uint8_t *)Sv;
	uint64_t cap = 1024, n = 0;
	uint64_t *pairs = (uint64_t *)malloc(cap * 16);
	if (!pairs) return -3;
	for (uint64_t j = 0; j < ns; j++) {
		const uint64_t sk = key_at(S, j, W);
		for (uint64_t i = 0; i < nr; i++) {
			if (key_at(R, i, W) != sk) continue;
			if (n == cap) {
				uint64_t *p = (uint64_t *)realloc(pairs, cap * 32);
				if (!p) {
					free(pairs);
					return -3;
				}
				pairs = p;
				cap *= 2;
			}
			pairs[2 * n] = rid_at(R, i, W);
			pairs[2 * n + 1] = rid_at(S, j, W);
			n++;
			if (lp) break;
		}
	}
	uint64_t sum = 0;
	const int rc = or_ptr_retrieve(pairs, n, Rv, nr, Sv, ns, W, &sum);
	free(pairs);
	if (rc) return rc;
	out->matched_cnt = n;
	out->check_sum = sum;
	out->buildpart_time = out->probejoin_time = out->total_time = 0;
	return 0;
}

/* The restated joins on the heads.  The retrieve sum is additive over pairs,
 * so it equals the plain join checksum of heads whose row_id is replaced by
 * the row id it points at: {key, rel[row_id - 1].row_id}.  A second join of
 * heads carrying 1 for an out-of-range row id (else 0) sums the pairs'
 * out-of-range ends, so a non-zero sum is the reference's out-of-bounds read.
 * algo: 0 nphj_sc, 1 nphj_lp, 2 phj_radix with per-partition `kind`. */
static int ptr_heads(const uint8_t *rel, uint64_t n, uint32_t W, int flag_only, uint64_t *out) {
	for (uint64_t i = 0; i < n; i++) {
		const uint64_t r = rid_at(rel, i, W);
		const int ok = r - 1 < n;
		out[2 * i] = key_at(rel, i, W);
		out[2 * i + 1] = flag_only ? (uint64_t)!ok : (ok ? rid_at(rel, r - 1, W) : 0);
	}
	return 0;
}

static int ptr_inner(int algo, int kind, uint32_t bits, uint32_t passes, const uint64_t *r, uint64_t nr,
                     const uint64_t *s, uint64_t ns, int threads, or_result_t *out) {
	if (algo == 0) return or_nphj_sc(r, nr, s, ns, 16, threads, out);
	if (algo == 1) return or_nphj_lp(r, nr, s, ns, 16, threads, out);
	return or_phj_radix(r, nr, s, ns, 16, bits, passes, kind, threads, out);
}

int or_ptr_join(int algo, int kind, uint32_t radix_bits, uint32_t passes, const void *Rv, uint64_t nr,
                const void *Sv, uint64_t ns, uint32_t W, int threads, or_result_t *out) {
	uint64_t *r = (uint64_t *)malloc((nr ? nr : 1) * 16), *s = (uint64_t *)malloc((ns ? ns : 1) * 16);
	if (!r || !s) {
		free(r);
		free(s);
		return -3;
	}
	or_result_t flag;
	ptr_heads((const uint8_t *)Rv, nr, W, 1, r);
	ptr_heads((const uint8_t *)Sv, ns, W, 1, s);
	int rc = ptr_inner(algo, kind, radix_bits, passes, r, nr, s, ns, threads, &flag);
	if (!rc && flag.check_sum) rc = -1;
	if (!rc) {
		ptr_heads((const uint8_t *)Rv, nr, W, 0, r);
		ptr_heads((const uint8_t *)Sv, ns, W, 0, s);
		rc = ptr_inner(algo, kind, radix_bits, passes, r, nr, s, ns, threads, out);
	}
	free(r);
	free(s);
	return rc;
}

/* ---- the reference's generator (src/gen.c, src/inc/eth_genzipf.c) --------
 * Restated on this C library's own srand()/rand(), the generator the
 * reference calls, so the draws are glibc's by construction.  kind: 0 pkfk
 * (gen_pkfk_uniform :192-214 + gen_fk_from_pk :174-189), 1 dupfk
 * (gen_dup_uniform :286-331), 2 zipf (pk_AR + gen_fk_zipfian_eth :361-377),
 * 3 payload (payload_gen_pkfk_uniform :246-283).  exact: also the
 * knuth_shuffle calls (:161-171).  Tuple bytes past the row id are zero.
 * Zipf table lookups run on `threads` threads (the draws stay sequential). */
static void ref_put(uint8_t *b, uint64_t i, uint32_t W, uint64_t key, uint64_t rid) {
	uint64_t *t = (uint64_t *)(b + i * W);
	t[0] = key;
	t[1] = rid;
	memset(t + 2, 0, W - 16);
}

static void ref_shuffle(uint8_t *b, uint64_t n, uint32_t W) {
	uint8_t tmp[4096];
	for (uint64_t i = 0; i < n; i++) {
		const uint64_t j = (uint64_t)rand() % n;
		memcpy(tmp, b + j * W, W);
		memcpy(b + j * W, b + i * W, W);
		memcpy(b + i * W, tmp, W);
	}
}

typedef struct {
	const int *draws;
	const double *lut;
	const uint64_t *alpha;
	uint64_t n, lo, hi;
	uint8_t *S;
	uint32_t W;
} zipf_arg_t;

static void *zipf_thread(void *p) {
	zipf_arg_t *a = (zipf_arg_t *)p;
	for (uint64_t j = a->lo; j < a->hi; j++) {
		const double r = (double)a->draws[j] / RAND_MAX;
		uint64_t pos = 0;
		if (a->lut[0] < r) {
			uint64_t l = 1, h = a->n;
			while (h - l > 1) {
				const uint64_t m = (l + h) / 2;
				if (a->lut[m] < r)
					l = m;
				else
					h = m;
			}
			pos = h;
		}
		ref_put(a->S, j, a->W, a->alpha[pos], j + 1);
	}
	return NULL;
}

int or_ref_generate(int kind, uint64_t nr, uint64_t ns, uint32_t W, double zf, int seed, int exact, int threads,
                    void *Rv, void *Sv) {
	uint8_t *R = (uint8_t *)Rv, *S = (uint8_t *)Sv;
	if (W < 16 || W % 16 || W > 4096 || nr == 0) return -2;
	if (kind != 2 && (ns % nr || (nr * W) % 4096)) return -2;
	srand((unsigned)seed);
	if (kind == 1) {
		int u = rand() % RAND_MAX;
		for (uint64_t i = 0; i < nr; i++) {
			const int k = u / 127773;
			u = 16807 * (u - k * 127773) - k * 2836;
			if (u < 0) u += 2147483647;
			const double r = (double)u * 4.656612875E-10;
			volatile double lo = (1.0 - r) * (1.0 - 0.5), hi = r * ((double)nr + 0.5);
			long long key = (int)round(lo + hi);
			if (key < 1) key = 1;
			if (key > (long long)nr) key = (long long)nr;
			ref_put(R, i, W, (uint64_t)key, i + 1);
		}
	} else {
		for (uint64_t i = 0; i < nr; i++) ref_put(R, i, W, i + 1, i + 1);
	}
	if (exact) ref_shuffle(R, nr, W);
	if (kind != 2) {
		for (uint64_t c = 0; c < ns / nr; c++) memcpy(S + c * nr * W, R, nr * W);
		if (exact) ref_shuffle(S, ns, W);
		if (kind == 3) {
			for (uint64_t j = 0; j < ns; j++) ((uint64_t *)(S + j * W))[1] = j + 1;
			for (uint64_t i = 0; i < nr; i++) ((uint64_t *)(R + i * W))[1] = i + 1;
		}
		return 0;
	}
	/* zipf: a fresh srand(seed), then alphabet, table and stream */
	if (!ns) return 0;
	srand((unsigned)seed);
	uint64_t *alpha = (uint64_t *)malloc(nr * 8);
	double *lut = (double *)malloc(nr * 8);
	int *draws = (int *)malloc(ns * sizeof(int));
	if (!alpha || !lut || !draws) {
		free(alpha);
		free(lut);
		free(draws);
		return -3;
	}
	for (uint64_t i = 0; i < nr; i++) alpha[i] = i + 1;
	for (uint64_t i = nr - 1; i > 0; i--) {
		const uint64_t k = i * (uint64_t)rand() / RAND_MAX;
		const uint64_t t = alpha[i];
		alpha[i] = alpha[k];
		alpha[k] = t;
	}
	double scale = 0, sum = 0;
	for (uint64_t i = 1; i <= nr; i++) scale += 1.0 / pow((double)i, zf);
	for (uint64_t i = 1; i <= nr; i++) {
		sum += 1.0 / pow((double)i, zf);
		lut[i - 1] = sum / scale;
	}
	for (uint64_t j = 0; j < ns; j++) draws[j] = rand();
	if (threads < 1) threads = 1;
	pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * threads);
	zipf_arg_t *args = (zipf_arg_t *)malloc(sizeof(zipf_arg_t) * threads);
	const uint64_t per = (ns + threads - 1) / threads;
	for (int t = 0; t < threads; t++) {
		const uint64_t lo = (uint64_t)t * per < ns ? (uint64_t)t * per : ns;
		const uint64_t hi = lo + per < ns ? lo + per : ns;
		args[t] = (zipf_arg_t){draws, lut, alpha, nr, lo, hi, S, W};
		pthread_create(&th[t], NULL, zipf_thread, &args[t]);
	}
	for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
	free(th);
	free(args);
	free(alpha);
	free(lut);
	free(draws);
	return 0;
}
