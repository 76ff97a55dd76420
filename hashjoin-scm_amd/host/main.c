/*
 * main.c -- command-line driver, the drop-in for the reference's bin/main
 * (src/main.c:147-550).
 *
 *   hjmain --algo=<name> --workload=<w> --subtype=<s> --param=<p> [options]
 *
 * The four reference arguments keep their meaning: the algorithm is looked
 * up by name in the library's registry (src/main.c:204-218, unknown name ->
 * error exit), and (workload, param, subtype) name a workload of the
 * reference's config tree (src/main.c:244-255, config/data/<w>/<p>_<s>.cfg).
 * The relations are generated on the GPU with the same layout and
 * closed-form results as the reference's generator; the dupfk and Zipf
 * workloads, whose results depend on glibc's draws, come from the library's
 * restatement of the reference generator (hj_ref_generate) on the host, as
 * does any reference workload with --exact (the reference's tuple order
 * too).  --cfg=FILE --data-prefix=DIR instead reads a reference data
 * configuration and its raw .bin files (src/algo/phj.c:1885-1915), and
 * --write-bin=PREFIX writes the relations as PREFIX_R.bin / PREFIX_S.bin.
 * The trial protocol is src/main.c:429-459: TRIALS runs, each printing
 * matched_cnt/check_sum and the phase times, then the mean of trials 2..N.
 *
 * Extra options (all optional):
 *   --r=N --s=N --width=W --gen=pkfk|pkfk_ref|payload_ref|zipf|random
 *   --theta=T --key-range=K --seed=S --bits=B --passes=P --bits-pass1=B1 --trials=T
 *   --device=D --list --cfg=FILE --data-prefix=DIR --write-bin=PREFIX --exact
 * Workload "baseline" with subtype C1, C2, C3 selects the benchmark
 * configurations: C1 2^20 x 2^24 uniform PK-FK, C2 2^24 x 2^28, C3 64-byte
 * Zipf 2^24 x 2^27 (theta from --theta, default 1.0).
 *
 * TPC-H, the drop-in for the reference's bin/tpch_query
 * (src/tpch/tpch_query.c:1129-1190, which fixes the query at build time):
 *   hjmain --tpch=14|19 --algo=tpch_rdx_bc --sf=1|10|100
 *          [--tbl-dir=DIR | --tpch-bin-dir=DIR] [--seed=S] [--trials=T]
 * reads DIR/part.tbl and DIR/lineitem.tbl (dbgen) or DIR/part.bin and
 * DIR/lineitem.bin (the reference's tbl2bin output), or generates the
 * scale factor's cardinalities on the GPU, and prints the reference's
 * result line (tpch_query.c:739-745) and phase times per trial.
 * --tbl2bin=DIR converts DIR/part.tbl and DIR/lineitem.tbl into the
 * reference's .bin files next to them, as bin/tpch_tbl2bin does.
 */
#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hjmi.h"

typedef struct {
	const char *workload, *param, *subtype;
	uint64_t r_num, s_num;
	uint32_t width, gen;
	const char *ref;  /* reference workload_name: host restatement of gen.c */
	int ref_only;     /* no closed-form device generator for it */
} wl_t;

/* Sizes from the reference's config/data and config/fndly-data files. */
static const wl_t k_workloads[] = {
    {"uniform", "pkfk", "A", 1ull << 24, 1ull << 28, 16, HJ_WL_PKFK_REF, "uniform_pkfk_A", 0},
    {"uniform", "pkfk", "B", 1ull << 26, 1ull << 28, 16, HJ_WL_PKFK_REF, "uniform_pkfk_B", 0},
    {"uniform", "pkfk", "C", 1ull << 28, 1ull << 28, 16, HJ_WL_PKFK_REF, "uniform_pkfk_C", 0},
    {"uniform", "dupfk", "D", 1ull << 24, 1ull << 28, 16, 0, "uniform_dupfk_D", 1},
    {"uniform", "dupfk", "E", 1ull << 26, 1ull << 28, 16, 0, "uniform_dupfk_E", 1},
    {"uniform", "dupfk", "F", 1ull << 28, 1ull << 28, 16, 0, "uniform_dupfk_F", 1},
    {"zipfian", "fk", "1.050_A", 1ull << 24, 1ull << 28, 16, 0, "zipfian_1.050_A", 1},
    {"zipfian", "fk", "1.250_A", 1ull << 24, 1ull << 28, 16, 0, "zipfian_1.250_A", 1},
    {"zipfian", "fk", "1.500_A", 1ull << 24, 1ull << 28, 16, 0, "zipfian_1.500_A", 1},
    {"zipfian", "fk", "1.750_A", 1ull << 24, 1ull << 28, 16, 0, "zipfian_1.750_A", 1},
    {"payload", "pkfk", "16", 1ull << 24, 1ull << 28, 16, HJ_WL_PAYLOAD_REF, "payload_pkfk_16", 0},
    {"payload", "pkfk", "32", 1ull << 24, 1ull << 28, 32, HJ_WL_PAYLOAD_REF, "payload_pkfk_32", 0},
    {"payload", "pkfk", "64", 1ull << 24, 1ull << 28, 64, HJ_WL_PAYLOAD_REF, "payload_pkfk_64", 0},
    {"payload", "pkfk", "128", 1ull << 24, 1ull << 28, 128, HJ_WL_PAYLOAD_REF, "payload_pkfk_128", 0},
    {"payload", "pkfk", "256", 1ull << 24, 1ull << 28, 256, HJ_WL_PAYLOAD_REF, "payload_pkfk_256", 0},
    {"payload", "pkfk", "512", 1ull << 24, 1ull << 28, 512, HJ_WL_PAYLOAD_REF, "payload_pkfk_512", 0},  /* ptr_* joins only */
    {"fndly", "pkfk", "Q", 1ull << 28, 1ull << 30, 16, HJ_WL_PKFK_REF, NULL, 0},
    {"fndly", "pkfk", "R", 1ull << 28, 1ull << 31, 16, HJ_WL_PKFK_REF, NULL, 0},
    {"fndly", "pkfk", "H", 1ull << 28, 1ull << 32, 16, HJ_WL_PKFK_REF, NULL, 0},
    {"fndly", "pkfk", "T", 1ull << 29, 1ull << 32, 16, HJ_WL_PKFK_REF, NULL, 0},
    {"baseline", "pkfk", "C1", 1ull << 20, 1ull << 24, 16, HJ_WL_PKFK_UNIFORM, NULL, 0},
    {"baseline", "pkfk", "C2", 1ull << 24, 1ull << 28, 16, HJ_WL_PKFK_UNIFORM, NULL, 0},
    {"baseline", "zipf", "C3", 1ull << 24, 1ull << 27, 64, HJ_WL_ZIPF, NULL, 0},
    {NULL, NULL, NULL, 0, 0, 0, 0, NULL, 0},
};

static int die(hj_ctx_t *ctx, int rc, const char *what) {
	fprintf(stderr, "[ERROR] %s: %s (%s)\n", what, hj_strerror(rc), ctx ? hj_last_error(ctx) : "");
	return 1;
}

/* ---- TPC-H ---------------------------------------------------------------- */
#define TPCH_PART_COLS 5
#define TPCH_LINE_COLS 7
static const uint64_t k_part_w[TPCH_PART_COLS] = {8, 10, 25, 4, 10};
static const uint64_t k_line_w[TPCH_LINE_COLS] = {8, 4, 8, 4, 10, 25, 10};

static void **part_cols(hj_tpch_part_t *p, int i) {
	void **c[TPCH_PART_COLS] = {(void **)&p->partkey, (void **)&p->brand, (void **)&p->type, (void **)&p->size,
	                            (void **)&p->container};
	return c[i];
}
static void **line_cols(hj_tpch_lineitem_t *l, int i) {
	void **c[TPCH_LINE_COLS] = {(void **)&l->partkey,  (void **)&l->quantity,     (void **)&l->extendedprice,
	                            (void **)&l->discount, (void **)&l->shipdate,     (void **)&l->shipinstruct,
	                            (void **)&l->shipmode};
	return c[i];
}

static int run_tpch(int query, const char *algo, int sf, const char *tbl_dir, const char *bin_dir, uint64_t seed,
                    int trials, int device) {
	int found = 0;
	for (int i = 0; i < hj_tpch_algo_count(); i++)
		if (!strcmp(hj_tpch_algo_name(i), algo)) found = 1;
	if (!found) {
		printf("[ERROR] Join algorithm named `%s' does not exist!\n", algo);
		return 1;
	}
	hj_tpch_part_t hp, dp;
	hj_tpch_lineitem_t hl, dl;
	memset(&hp, 0, sizeof hp), memset(&dp, 0, sizeof dp), memset(&hl, 0, sizeof hl), memset(&dl, 0, sizeof dl);
	char pp[1024], lp[1024];
	int rc;
	if (tbl_dir || bin_dir) {
		const char *dir = tbl_dir ? tbl_dir : bin_dir;
		snprintf(pp, sizeof pp, "%s/part.%s", dir, tbl_dir ? "tbl" : "bin");
		snprintf(lp, sizeof lp, "%s/lineitem.%s", dir, tbl_dir ? "tbl" : "bin");
		if (tbl_dir) {
			if (hj_tpch_tbl_rows(pp, &hp.rows) || hj_tpch_tbl_rows(lp, &hl.rows)) {
				fprintf(stderr, "[ERROR] cannot read %s or %s\n", pp, lp);
				return 1;
			}
		} else {
			FILE *f = fopen(pp, "rb"), *g = fopen(lp, "rb");
			if (!f || !g || fseek(f, 0, SEEK_END) || fseek(g, 0, SEEK_END)) {
				fprintf(stderr, "[ERROR] cannot read %s or %s\n", pp, lp);
				return 1;
			}
			hp.rows = (uint64_t)ftell(f) / 168, hl.rows = (uint64_t)ftell(g) / 175;
			fclose(f), fclose(g);
		}
	} else if ((rc = hj_tpch_cardinality(sf, &hp.rows, &hl.rows))) {
		fprintf(stderr, "[ERROR] unsupported scaling factor %d\n", sf);
		return 1;
	}
	dp.rows = hp.rows, dl.rows = hl.rows;
	hj_ctx_t *ctx = NULL;
	if ((rc = hj_ctx_create(device, &ctx))) return die(NULL, rc, "hj_ctx_create");
	for (int i = 0; i < TPCH_PART_COLS; i++)
		if ((rc = hj_dev_alloc(ctx, dp.rows * k_part_w[i] + 16, part_cols(&dp, i)))) return die(ctx, rc, "alloc part");
	for (int i = 0; i < TPCH_LINE_COLS; i++)
		if ((rc = hj_dev_alloc(ctx, dl.rows * k_line_w[i] + 16, line_cols(&dl, i)))) return die(ctx, rc, "alloc lineitem");
	if (tbl_dir || bin_dir) {
		for (int i = 0; i < TPCH_PART_COLS; i++)
			if (!(*part_cols(&hp, i) = malloc(hp.rows * k_part_w[i] + 16))) return die(ctx, HJ_ENOMEM, "host part");
		for (int i = 0; i < TPCH_LINE_COLS; i++)
			if (!(*line_cols(&hl, i) = malloc(hl.rows * k_line_w[i] + 16))) return die(ctx, HJ_ENOMEM, "host lineitem");
		rc = tbl_dir ? hj_tpch_tbl_read_part(pp, &hp) : hj_tpch_bin_read_part(pp, &hp);
		if (rc) return die(ctx, rc, pp);
		rc = tbl_dir ? hj_tpch_tbl_read_lineitem(lp, &hl) : hj_tpch_bin_read_lineitem(lp, &hl);
		if (rc) return die(ctx, rc, lp);
		for (int i = 0; i < TPCH_PART_COLS; i++)
			if ((rc = hj_copy_h2d(ctx, *part_cols(&dp, i), *part_cols(&hp, i), hp.rows * k_part_w[i])))
				return die(ctx, rc, "copy part");
		for (int i = 0; i < TPCH_LINE_COLS; i++)
			if ((rc = hj_copy_h2d(ctx, *line_cols(&dl, i), *line_cols(&hl, i), hl.rows * k_line_w[i])))
				return die(ctx, rc, "copy lineitem");
		for (int i = 0; i < TPCH_PART_COLS; i++) free(*part_cols(&hp, i));
		for (int i = 0; i < TPCH_LINE_COLS; i++) free(*line_cols(&hl, i));
	} else if ((rc = hj_tpch_generate(ctx, seed, &dp, &dl))) {
		return die(ctx, rc, "hj_tpch_generate");
	}
	printf("CONFIGURATION:\tQUERY: %d\tSF: %d\tpart_num: %llu\tlineitem_num: %llu\n", query, sf,
	       (unsigned long long)dp.rows, (unsigned long long)dl.rows);
	for (int t = 0; t < trials; t++) {
		hj_tpch_result_t res;
		if ((rc = hj_tpch_query(ctx, query, algo, &dp, &dl, &res))) return die(ctx, rc, "hj_tpch_query");
		printf("matched_cnt: %llu\t", (unsigned long long)res.matched_cnt);
		if (query == 14)
			printf("promo_revenue: %.2f\tnumerator: %.2f\tdenominator: %.2f\n", res.promo_revenue, res.numerator,
			       res.denominator);
		else
			printf("revenue: %.2f\n", res.revenue);
		printf("algo: %s\tpart_selected: %llu\tlineitem_selected: %llu\tpushdown_time: %.9f\tbuildpart_time: "
		       "%.9f\tprobejoin_time: %.9f\ttotal_time: %.9f\n",
		       algo, (unsigned long long)res.part_selected, (unsigned long long)res.lineitem_selected,
		       res.pushdown_time, res.buildpart_time, res.probejoin_time, res.total_time);
	}
	for (int i = 0; i < TPCH_PART_COLS; i++) hj_dev_free(ctx, *part_cols(&dp, i));
	for (int i = 0; i < TPCH_LINE_COLS; i++) hj_dev_free(ctx, *line_cols(&dl, i));
	hj_ctx_destroy(ctx);
	return 0;
}

/* host tuples of a reference workload, copied into device memory */
static int ref_upload(hj_ctx_t *ctx, const char *name, uint32_t width, int exact, void *r_dev, void *s_dev,
                      uint64_t *nr, uint64_t *ns) {
	hj_ref_spec_t sp;
	int rc = hj_ref_spec_for(name, width, &sp);
	if (rc) return rc;
	sp.exact_order = (uint32_t)exact;
	if (*nr) sp.r_num = *nr;
	if (*ns) sp.s_num = *ns;
	void *r = malloc(sp.r_num * sp.tuple_size), *s = malloc(sp.s_num * sp.tuple_size + 1);
	if (!r || !s) {
		free(r);
		free(s);
		return HJ_ENOMEM;
	}
	rc = hj_ref_generate(&sp, r, s);
	if (!rc) rc = hj_copy_h2d(ctx, r_dev, r, sp.r_num * sp.tuple_size);
	if (!rc) rc = hj_copy_h2d(ctx, s_dev, s, sp.s_num * sp.tuple_size);
	free(r);
	free(s);
	*nr = sp.r_num;
	*ns = sp.s_num;
	return rc;
}

int main(int argc, char **argv) {
	const char *algo = NULL, *workload = "baseline", *subtype = "C2", *param = "pkfk", *gen = NULL;
	const char *cfg = NULL, *data_prefix = ".", *write_bin = NULL;
	int exact = 0, tpch = 0, sf = 1;
	const char *tbl_dir = NULL, *tpch_bin_dir = NULL, *tbl2bin_dir = NULL;
	uint64_t r_num = 0, s_num = 0, key_range = 0, seed = 12345;
	uint32_t width = 0, bits = 0, passes = 0, bits_pass1 = 0;
	int trials = 6, device = 0, list = 0;
	double theta = 1.0;
	static struct option opts[] = {
	    {"algo", required_argument, 0, 'a'},    {"workload", required_argument, 0, 'w'},
	    {"subtype", required_argument, 0, 't'}, {"param", required_argument, 0, 'p'},
	    {"r", required_argument, 0, 'r'},       {"s", required_argument, 0, 's'},
	    {"width", required_argument, 0, 'W'},   {"gen", required_argument, 0, 'g'},
	    {"theta", required_argument, 0, 'z'},   {"key-range", required_argument, 0, 'k'},
	    {"seed", required_argument, 0, 'S'},    {"bits", required_argument, 0, 'b'},
	    {"passes", required_argument, 0, 'P'},  {"trials", required_argument, 0, 'T'},
	    {"device", required_argument, 0, 'd'},  {"list", no_argument, 0, 'l'},
	    {"cfg", required_argument, 0, 'c'},     {"data-prefix", required_argument, 0, 'D'},
	    {"write-bin", required_argument, 0, 'o'}, {"exact", no_argument, 0, 'x'},
	    {"bits-pass1", required_argument, 0, '1'}, {"tpch", required_argument, 0, 'Q'},
	    {"sf", required_argument, 0, 'F'},      {"tbl-dir", required_argument, 0, 'L'},
	    {"tpch-bin-dir", required_argument, 0, 'B'}, {"tbl2bin", required_argument, 0, 'C'},
	    {0, 0, 0, 0}};
	int c;
	while ((c = getopt_long(argc, argv, "", opts, NULL)) != -1) {
		switch (c) {
		case 'a': algo = optarg; break;
		case 'w': workload = optarg; break;
		case 't': subtype = optarg; break;
		case 'p': param = optarg; break;
		case 'r': r_num = strtoull(optarg, NULL, 0); break;
		case 's': s_num = strtoull(optarg, NULL, 0); break;
		case 'W': width = (uint32_t)strtoul(optarg, NULL, 0); break;
		case 'g': gen = optarg; break;
		case 'z': theta = atof(optarg); break;
		case 'k': key_range = strtoull(optarg, NULL, 0); break;
		case 'S': seed = strtoull(optarg, NULL, 0); break;
		case 'b': bits = (uint32_t)strtoul(optarg, NULL, 0); break;
		case 'P': passes = (uint32_t)strtoul(optarg, NULL, 0); break;
		case 'T': trials = atoi(optarg); break;
		case 'd': device = atoi(optarg); break;
		case 'l': list = 1; break;
		case 'c': cfg = optarg; break;
		case 'D': data_prefix = optarg; break;
		case 'o': write_bin = optarg; break;
		case 'x': exact = 1; break;
		case '1': bits_pass1 = (uint32_t)strtoul(optarg, NULL, 0); break;
		case 'Q': tpch = atoi(optarg); break;
		case 'F': sf = atoi(optarg); break;
		case 'L': tbl_dir = optarg; break;
		case 'B': tpch_bin_dir = optarg; break;
		case 'C': tbl2bin_dir = optarg; break;
		default: fprintf(stderr, "unknown option\n"); return 2;
		}
	}
	if (tbl2bin_dir) {
		const char *tables[2] = {"part", "lineitem"};
		for (int i = 0; i < 2; i++) {
			char in[1024], out[1024];
			uint64_t rows = 0;
			snprintf(in, sizeof in, "%s/%s.tbl", tbl2bin_dir, tables[i]);
			snprintf(out, sizeof out, "%s/%s.bin", tbl2bin_dir, tables[i]);
			if (hj_tpch_tbl2bin(tables[i], in, out, &rows)) {
				fprintf(stderr, "[ERROR] tbl2bin %s\n", in);
				return 1;
			}
			printf("%s_num: %llu\n", tables[i], (unsigned long long)rows);
		}
		return 0;
	}
	if (tpch) {
		if (tpch != 14 && tpch != 19) {
			fprintf(stderr, "[ERROR] --tpch must be 14 or 19\n");
			return 2;
		}
		return run_tpch(tpch, algo ? algo : "tpch_nphj_sc", sf, tbl_dir, tpch_bin_dir, seed, trials == 6 ? 1 : trials,
		                device);
	}
	if (list) {
		for (int i = 0; i < hj_algo_count(); i++) printf("%s\n", hj_algo_name(i));
		return 0;
	}
	if (!algo) {
		fprintf(stderr, "[ERROR] --algo is required\n");
		return 2;
	}
	int found = 0;
	for (int i = 0; i < hj_algo_count(); i++)
		if (strcmp(hj_algo_name(i), algo) == 0) found = 1;
	if (!found) {
		printf("[ERROR] Join algorithm named `%s' does not exist!\n", algo);
		return 1;
	}
	hj_gen_spec_t spec;
	memset(&spec, 0, sizeof spec);
	spec.seed = seed;
	spec.theta = theta;
	const wl_t *wl = NULL;
	hj_data_cfg_t dc;
	char wname[256];
	if (cfg) {
		int rc = hj_data_cfg_read(cfg, &dc);
		if (rc || dc.r_tuple_size != dc.s_tuple_size || !dc.r_tuple_size) {
			fprintf(stderr, "[ERROR] cannot use data configuration %s\n", cfg);
			return 1;
		}
		spec.r_num = dc.r_tuple_num;
		spec.s_num = dc.s_tuple_num;
		spec.width = dc.r_tuple_size;
		snprintf(wname, sizeof wname, "%s", dc.workload_name);
	} else {
		snprintf(wname, sizeof wname, "%s_%s_%s", workload, param, subtype);
	}
	for (const wl_t *w = k_workloads; !cfg && w->workload; w++)
		if (!strcmp(w->workload, workload) && !strcmp(w->param, param) && !strcmp(w->subtype, subtype)) wl = w;
	if (cfg) {
	} else if (wl) {
		spec.r_num = wl->r_num;
		spec.s_num = wl->s_num;
		spec.width = wl->width;
		spec.workload = wl->gen;
	} else if (!r_num || !s_num) {
		fprintf(stderr, "[ERROR] unknown workload %s/%s_%s and no --r/--s given\n", workload, param, subtype);
		return 1;
	} else {
		spec.width = 16;
		spec.workload = HJ_WL_PKFK_UNIFORM;
	}
	if (r_num && !cfg) spec.r_num = r_num;
	if (s_num && !cfg) spec.s_num = s_num;
	if (width && !cfg) spec.width = width;
	if (gen) {
		if (!strcmp(gen, "pkfk")) spec.workload = HJ_WL_PKFK_UNIFORM;
		else if (!strcmp(gen, "pkfk_ref")) spec.workload = HJ_WL_PKFK_REF;
		else if (!strcmp(gen, "payload_ref")) spec.workload = HJ_WL_PAYLOAD_REF;
		else if (!strcmp(gen, "zipf")) spec.workload = HJ_WL_ZIPF;
		else if (!strcmp(gen, "random")) spec.workload = HJ_WL_RANDOM;
		else {
			fprintf(stderr, "[ERROR] unknown --gen %s\n", gen);
			return 1;
		}
	}
	spec.key_range = key_range ? key_range : spec.r_num;

	hj_ctx_t *ctx = NULL;
	int rc = hj_ctx_create(device, &ctx);
	if (rc) return die(NULL, rc, "hj_ctx_create");
	void *r_dev = NULL, *s_dev = NULL;
	if ((rc = hj_dev_alloc(ctx, spec.r_num * spec.width, &r_dev))) return die(ctx, rc, "alloc R");
	if ((rc = hj_dev_alloc(ctx, spec.s_num * spec.width, &s_dev))) return die(ctx, rc, "alloc S");
	if (cfg) {
		char path[1024];
		snprintf(path, sizeof path, "%s/%s", data_prefix, dc.r_path_suffix);
		if ((rc = hj_bin_read(ctx, path, spec.r_num, spec.width, r_dev))) return die(ctx, rc, "read R");
		snprintf(path, sizeof path, "%s/%s", data_prefix, dc.s_path_suffix);
		if ((rc = hj_bin_read(ctx, path, spec.s_num, spec.width, s_dev))) return die(ctx, rc, "read S");
	} else if (wl && wl->ref && (wl->ref_only || exact)) {
		uint64_t nr = r_num, ns = s_num;
		if ((rc = ref_upload(ctx, wl->ref, spec.width, exact, r_dev, s_dev, &nr, &ns)))
			return die(ctx, rc, "reference generator");
	} else if ((rc = hj_generate(ctx, &spec, r_dev, s_dev))) {
		return die(ctx, rc, "generate");
	}
	if (write_bin) {
		char path[1024];
		snprintf(path, sizeof path, "%s_R.bin", write_bin);
		if ((rc = hj_bin_write(ctx, path, r_dev, spec.r_num, spec.width))) return die(ctx, rc, "write R");
		snprintf(path, sizeof path, "%s_S.bin", write_bin);
		if ((rc = hj_bin_write(ctx, path, s_dev, spec.s_num, spec.width))) return die(ctx, rc, "write S");
	}

	printf("CONFIGURATION:\tTUPLE_T_SIZE: %u\tNUM_RADIX_BITS: %u\tNUM_PASSES: %u\tTRIALS: %d\n", spec.width, bits,
	       passes, trials);
	hj_relation_t R = {r_dev, spec.r_num, spec.width, HJ_LOC_DEVICE};
	hj_relation_t S = {s_dev, spec.s_num, spec.width, HJ_LOC_DEVICE};
	hj_params_t prm;
	memset(&prm, 0, sizeof prm);
	prm.radix_bits = bits;
	prm.passes = passes;
	prm.radix_bits_pass1 = bits_pass1;
	double sum_b = 0, sum_p = 0, sum_t = 0;
	int counted = 0;
	for (int t = 0; t < trials; t++) {
		hj_result_t res;
		if ((rc = hj_join(ctx, algo, &R, &S, &prm, &res))) return die(ctx, rc, "hj_join");
		printf("matched_cnt: %llu\tcheck_sum: %llu\n", (unsigned long long)res.matched_cnt,
		       (unsigned long long)res.check_sum);
		printf("memory_type: hbm\talgo: %s\tworkload: %s\tbuildpart_time: %.9f\tprobejoin_time: %.9f\ttotal_time: "
		       "%.9f\tradix_bits: %u\tpasses: %u\ttuples_per_s: %.4e\thbm_GBps: %.1f\n",
		       algo, wname, res.buildpart_time, res.probejoin_time, res.total_time, res.radix_bits, res.passes,
		       (double)(spec.r_num + spec.s_num) / res.total_time, res.algo_bytes / res.total_time * 1e-9);
		if (t > 0 || trials == 1) {
			sum_b += res.buildpart_time;
			sum_p += res.probejoin_time;
			sum_t += res.total_time;
			counted++;
		}
	}
	printf("\nnum_trials: %d\tmemory_type: hbm\talgo: %s\tworkload: %s\tbuildpart_time: %.9f\tprobejoin_time: "
	       "%.9f\ttotal_time: %.9f\n",
	       counted, algo, wname, sum_b / counted, sum_p / counted, sum_t / counted);
	hj_dev_free(ctx, r_dev);
	hj_dev_free(ctx, s_dev);
	hj_ctx_destroy(ctx);
	return 0;
}
