// hj_refdata.hip -- the reference's own datasets: its generator, its data
// configuration files and its raw tuple files (host code).
//
// Generator (src/gen.c, src/inc/eth_genzipf.c).  The reference draws every
// value from glibc rand() after srand(12345) (gen.c:144), so its datasets
// are fixed by glibc's generator: random_r's TYPE_3 additive feedback
// generator, x[i] = x[i-3] + x[i-31] mod 2^32 over a state seeded by the
// 16807 multiplicative congruential generator, the first 310 outputs
// discarded, output x >> 1 (the published description of glibc's
// random()).  hj_glibc_rand restates it; the families of gen.c follow it
// draw for draw:
//   pkfk    gen_pkfk_uniform (gen.c:192-214) + gen_fk_from_pk (:174-189)
//   dupfk   gen_dup_uniform (gen.c:286-331): keys from the 16807 generator
//           seeded with rand() % RAND_MAX, mapped to [min_key, max_key]
//   zipf    pk_AR of pkfk A + gen_fk_zipfian_eth (gen.c:361-377) =
//           gen_alphabet / gen_zipf_lut / gen_zipf (eth_genzipf.c:30-148)
//   payload payload_gen_pkfk_uniform (gen.c:246-283)
// including knuth_shuffle (gen.c:161-171) when exact_order is set, so the
// key and row_id words equal the reference's files byte for byte (the rest
// of a tuple is uninitialised stack memory there and zero here).  Without
// exact_order the shuffles are skipped: the same tuples in another order,
// which gives every join the same result in a fraction of the time.
//
// Data configuration files (config/data/**/*.cfg, read by data_cfg,
// src/inc/utils.c:236-290): the libconfig subset they use -- `name = value;`
// settings with string, integer and float values and comments.
//
// Raw tuple files: packed tuple_t[] with no header (gen.c writes them with
// pmem_memcpy_persist); read and written through pinned staging buffers.

#include "hj_internal.h"

#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

namespace {

// glibc random() TYPE_3 (degree 31, separation 3) after srandom(seed)
struct hj_glibc_rand {
	uint32_t ring[31];  // ring[p] holds x[i - 31] for the next output x[i]
	int p = 0;

	explicit hj_glibc_rand(int32_t seed) {
		int32_t x[34];
		x[0] = seed == 0 ? 1 : seed;
		for (int i = 1; i < 31; i++) {
			// 16807 * x mod (2^31 - 1) without overflow (Schrage)
			const int64_t hi = x[i - 1] / 127773, lo = x[i - 1] % 127773;
			int64_t w = 16807 * lo - 2836 * hi;
			if (w < 0) w += 2147483647;
			x[i] = (int32_t)w;
		}
		for (int i = 31; i < 34; i++) x[i] = x[i - 31];
		for (int j = 0; j < 31; j++) ring[j] = (uint32_t)x[j + 3];
		for (int i = 0; i < 310; i++) (void)next();
	}
	int32_t next() {
		const uint32_t v = ring[p] + ring[(p + 28) % 31];
		ring[p] = v;
		p = p + 1 == 31 ? 0 : p + 1;
		return (int32_t)(v >> 1);
	}
};

const int32_t kRandMax = 2147483647;

// tuple_t: {u64 key; u64 row_id; zero rest}
inline void put_tuple(uint8_t *base, uint64_t i, uint32_t W, uint64_t key, uint64_t rid) {
	uint64_t *t = (uint64_t *)(base + i * W);
	t[0] = key;
	t[1] = rid;
	if (W > 16) memset(t + 2, 0, W - 16);
}
inline uint64_t *tuple_at(uint8_t *base, uint64_t i, uint32_t W) { return (uint64_t *)(base + i * W); }

// knuth_shuffle (gen.c:161-171): swap tuple idx with tuple rand() % n
void knuth_shuffle(hj_glibc_rand &g, uint8_t *base, uint64_t n, uint32_t W) {
	std::vector<uint8_t> tmp(W);
	for (uint64_t i = 0; i < n; i++) {
		const uint64_t j = (uint64_t)g.next() % n;
		memcpy(tmp.data(), base + j * W, W);
		memcpy(base + j * W, base + i * W, W);
		memcpy(base + i * W, tmp.data(), W);
	}
}

// gen_fk_from_pk (gen.c:174-189): ns / nr copies of R, then shuffled
void copies_of(hj_glibc_rand &g, const uint8_t *R, uint64_t nr, uint8_t *S, uint64_t ns, uint32_t W, bool exact) {
	for (uint64_t c = 0; c < ns / nr; c++) memcpy(S + c * nr * W, R, nr * W);
	if (exact) knuth_shuffle(g, S, ns, W);
}

// pk side of gen_pkfk_uniform / payload_gen_pkfk_uniform: key = row_id =
// 1..n, shuffled
void pk_side(hj_glibc_rand &g, uint8_t *R, uint64_t nr, uint32_t W, bool exact) {
	for (uint64_t i = 0; i < nr; i++) put_tuple(R, i, W, i + 1, i + 1);
	if (exact) knuth_shuffle(g, R, nr, W);
}

// Zipf probe stream of gen_zipf (eth_genzipf.c:98-148) over gen_alphabet
// (:30-51) and gen_zipf_lut (:58-90).  The draws are sequential; the table
// lookups run on `threads` threads.
void zipf_side(hj_glibc_rand &g, uint8_t *S, uint64_t ns, uint64_t alphabet_size, double zf, uint32_t W,
               unsigned threads) {
	std::vector<uint64_t> alphabet(alphabet_size);
	for (uint64_t i = 0; i < alphabet_size; i++) alphabet[i] = i + 1;
	for (uint64_t i = alphabet_size - 1; i > 0; i--) {
		const uint64_t k = i * (uint64_t)g.next() / (uint64_t)kRandMax;
		std::swap(alphabet[i], alphabet[k]);
	}
	std::vector<double> lut(alphabet_size);
	double scaling = 0.0, sum = 0.0;
	for (uint64_t i = 1; i <= alphabet_size; i++) scaling += 1.0 / pow((double)i, zf);
	for (uint64_t i = 1; i <= alphabet_size; i++) {
		sum += 1.0 / pow((double)i, zf);
		lut[i - 1] = sum / scaling;
	}
	std::vector<int32_t> draws(ns);
	for (uint64_t j = 0; j < ns; j++) draws[j] = g.next();
	auto work = [&](uint64_t lo, uint64_t hi) {
		for (uint64_t j = lo; j < hi; j++) {
			const double r = (double)draws[j] / kRandMax;
			uint64_t pos;
			if (lut[0] >= r) {
				pos = 0;
			} else {
				uint64_t left = 1, right = alphabet_size;
				while (right - left > 1) {
					const uint64_t m = (left + right) / 2;
					if (lut[m] < r)
						left = m;
					else
						right = m;
				}
				pos = right;
			}
			// pos < alphabet_size: lut[n-1] = 1 exactly (both sums add the
			// same terms in the same order) and r <= 1
			put_tuple(S, j, W, alphabet[std::min(pos, alphabet_size - 1)], j + 1);
		}
	};
	threads = std::max(1u, threads);
	std::vector<std::thread> pool;
	const uint64_t per = (ns + threads - 1) / threads;
	for (unsigned t = 0; t < threads; t++) {
		const uint64_t lo = std::min<uint64_t>(ns, t * per), hi = std::min<uint64_t>(ns, lo + per);
		pool.emplace_back(work, lo, hi);
	}
	for (auto &th : pool) th.join();
}

bool ends_with(const std::string &s, const char *suf) {
	const size_t n = strlen(suf);
	return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

}  // namespace

extern "C" {

int hj_ref_rand(int32_t seed, uint64_t n, int32_t *out) {
	if (n && !out) return HJ_EINVAL;
	hj_glibc_rand g(seed);
	for (uint64_t i = 0; i < n; i++) out[i] = g.next();
	return HJ_OK;
}

int hj_ref_spec_for(const char *name, uint32_t tuple_size, hj_ref_spec_t *out) {
	if (!name || !out) return HJ_EINVAL;
	memset(out, 0, sizeof *out);
	out->seed = 12345;
	out->tuple_size = tuple_size ? tuple_size : 16;
	out->s_num = 1ull << 28;
	const std::string n(name);
	if (n == "uniform_pkfk_A" || n == "uniform_pkfk_B" || n == "uniform_pkfk_C") {
		out->kind = HJ_REF_PKFK;
		out->r_num = 1ull << (n.back() == 'A' ? 24 : n.back() == 'B' ? 26 : 28);
	} else if (n == "uniform_dupfk_D" || n == "uniform_dupfk_E" || n == "uniform_dupfk_F") {
		out->kind = HJ_REF_DUPFK;
		out->r_num = 1ull << (n.back() == 'D' ? 24 : n.back() == 'E' ? 26 : 28);
	} else if (n.rfind("zipfian_", 0) == 0 && ends_with(n, "_A")) {
		out->kind = HJ_REF_ZIPF;
		out->r_num = 1ull << 24;
		out->zipf_factor = atof(n.c_str() + 8);  // "zipfian_1.050_A": 1 + cfg theta
		if (!(out->zipf_factor > 0)) return HJ_EINVAL;
	} else if (n.rfind("payload_pkfk_", 0) == 0) {
		out->kind = HJ_REF_PAYLOAD;
		out->r_num = 1ull << 24;
		out->tuple_size = (uint32_t)atoi(n.c_str() + 13);
	} else {
		return HJ_EINVAL;
	}
	return HJ_OK;
}

int hj_ref_generate(const hj_ref_spec_t *spec, void *r_host, void *s_host) {
	if (!spec) return HJ_EINVAL;
	const uint32_t W = spec->tuple_size;
	const uint64_t nr = spec->r_num, ns = spec->s_num;
	if (W < 16 || W % 16 || W > 4096) return HJ_EINVAL;
	if ((nr && !r_host) || (ns && !s_host) || nr == 0) return HJ_EINVAL;
	const bool exact = spec->exact_order != 0;
	uint8_t *R = (uint8_t *)r_host, *S = (uint8_t *)s_host;
	const unsigned threads = std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
	switch (spec->kind) {
	case HJ_REF_PKFK:
	case HJ_REF_PAYLOAD: {
		// gen_fk_from_pk copies whole 4 KiB blocks of R, |S| / |R| times
		if (ns % nr || (nr * W) % 4096) return HJ_EINVAL;
		hj_glibc_rand g(spec->seed);
		pk_side(g, R, nr, W, exact);
		copies_of(g, R, nr, S, ns, W, exact);
		if (spec->kind == HJ_REF_PAYLOAD) {
			for (uint64_t j = 0; j < ns; j++) tuple_at(S, j, W)[1] = j + 1;
			for (uint64_t i = 0; i < nr; i++) tuple_at(R, i, W)[1] = i + 1;
		}
		return HJ_OK;
	}
	case HJ_REF_DUPFK: {
		if (ns % nr || (nr * W) % 4096 || nr > (uint64_t)kRandMax) return HJ_EINVAL;
		hj_glibc_rand g(spec->seed);
		// the uniform generator gen.c:292-318 cites (Burkardt's uniform.c)
		int32_t u = g.next() % kRandMax;
		const int64_t min_key = 1, max_key = (int64_t)nr;
		for (uint64_t i = 0; i < nr; i++) {
			const int32_t k = u / 127773;
			u = 16807 * (u - k * 127773) - k * 2836;
			if (u < 0) u += kRandMax;
			double r = (double)u * 4.656612875E-10;
			// two roundings as written: (1-r)*(min-0.5) is exact for
			// min_key = 1, so a fused multiply-add of that product cannot
			// change the sum
			volatile double lo = (1.0 - r) * ((double)min_key - 0.5);
			volatile double hi = r * ((double)max_key + 0.5);
			r = lo + hi;
			int64_t key = (int)round(r);
			key = std::min(std::max(key, min_key), max_key);
			put_tuple(R, i, W, (uint64_t)key, i + 1);
		}
		if (exact) knuth_shuffle(g, R, nr, W);
		copies_of(g, R, nr, S, ns, W, exact);
		return HJ_OK;
	}
	case HJ_REF_ZIPF: {
		if (!(spec->zipf_factor > 0)) return HJ_EINVAL;
		{
			hj_glibc_rand g(spec->seed);  // pk_AR: the pk side of pkfk A
			pk_side(g, R, nr, W, exact);
		}
		if (ns) {
			hj_glibc_rand g(spec->seed);
			zipf_side(g, S, ns, nr, spec->zipf_factor, W, threads);
		}
		return HJ_OK;
	}
	default: return HJ_EINVAL;
	}
}

// ---------------------------------------------------------------- .cfg
int hj_data_cfg_read(const char *path, hj_data_cfg_t *out) {
	if (!path || !out) return HJ_EINVAL;
	memset(out, 0, sizeof *out);
	FILE *f = fopen(path, "rb");
	if (!f) return HJ_EINVAL;
	std::string txt;
	char buf[4096];
	size_t got;
	while ((got = fread(buf, 1, sizeof buf, f)) > 0) txt.append(buf, got);
	fclose(f);
	// strip comments: # ... and // ... to end of line, /* ... */
	std::string s;
	for (size_t i = 0; i < txt.size();) {
		if (txt[i] == '"') {  // strings are copied whole
			size_t j = txt.find('"', i + 1);
			if (j == std::string::npos) return HJ_EINVAL;
			s.append(txt, i, j - i + 1);
			i = j + 1;
		} else if (txt[i] == '#' || txt.compare(i, 2, "//") == 0) {
			while (i < txt.size() && txt[i] != '\n') i++;
		} else if (txt.compare(i, 2, "/*") == 0) {
			const size_t j = txt.find("*/", i + 2);
			if (j == std::string::npos) return HJ_EINVAL;
			i = j + 2;
		} else {
			s += txt[i++];
		}
	}
	size_t i = 0;
	auto skip_ws = [&]() {
		while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\r' || s[i] == '\n')) i++;
	};
	for (;;) {
		skip_ws();
		if (i >= s.size()) break;
		size_t j = i;
		while (j < s.size() && (isalnum((unsigned char)s[j]) || s[j] == '_' || s[j] == '-' || s[j] == '*')) j++;
		if (j == i) return HJ_EINVAL;
		const std::string name = s.substr(i, j - i);
		i = j;
		skip_ws();
		if (i >= s.size() || (s[i] != '=' && s[i] != ':')) return HJ_EINVAL;
		i++;
		skip_ws();
		std::string sval;
		double dval = 0;
		long long ival = 0;
		bool is_str = false;
		if (i < s.size() && s[i] == '"') {
			const size_t k = s.find('"', i + 1);
			sval = s.substr(i + 1, k - i - 1);
			i = k + 1;
			is_str = true;
		} else {
			size_t k = i;
			while (k < s.size() && s[k] != ';' && s[k] != ',' && s[k] != '\n') k++;
			std::string v = s.substr(i, k - i);
			while (!v.empty() && (v.back() == ' ' || v.back() == '\t' || v.back() == '\r')) v.pop_back();
			if (v.empty()) return HJ_EINVAL;
			char *e = nullptr;
			errno = 0;
			dval = strtod(v.c_str(), &e);
			ival = strtoll(v.c_str(), nullptr, 0);
			if (e == v.c_str()) return HJ_EINVAL;
			i = k;
		}
		skip_ws();
		if (i < s.size() && (s[i] == ';' || s[i] == ',')) i++;
		auto str = [&](char *dst, size_t cap) {
			snprintf(dst, cap, "%s", sval.c_str());
		};
		if (name == "r_path_suffix" && is_str) str(out->r_path_suffix, sizeof out->r_path_suffix);
		else if (name == "s_path_suffix" && is_str) str(out->s_path_suffix, sizeof out->s_path_suffix);
		else if (name == "workload_name" && is_str) str(out->workload_name, sizeof out->workload_name);
		else if (name == "r_tuple_num") out->r_tuple_num = (uint64_t)ival;
		else if (name == "s_tuple_num") out->s_tuple_num = (uint64_t)ival;
		else if (name == "r_tuple_size") out->r_tuple_size = (uint32_t)ival;
		else if (name == "s_tuple_size") out->s_tuple_size = (uint32_t)ival;
		else if (name == "min_key") out->min_key = ival;
		else if (name == "max_key") out->max_key = ival;
		else if (name == "theta") out->theta = dval;
		else if (name == "selectivity") out->selectivity = dval;
		else if (name == "density") out->density = dval;
		else if (name == "sparsity") out->sparsity = dval;
	}
	return HJ_OK;
}

// ---------------------------------------------------------------- .bin
static const size_t kStage = 64ull << 20;

int hj_bin_read(hj_ctx_t *ctx, const char *path, uint64_t tuple_num, uint32_t tuple_size, void *dev) {
	if (!ctx || !path || (tuple_num && !dev)) return HJ_EINVAL;
	const uint64_t bytes = tuple_num * tuple_size;
	FILE *f = fopen(path, "rb");
	if (!f) return hj_fail(ctx, HJ_EINVAL, "cannot open %s: %s", path, strerror(errno));
	if (fseeko(f, 0, SEEK_END) || (uint64_t)ftello(f) < bytes) {
		fclose(f);
		return hj_fail(ctx, HJ_EINVAL, "%s holds fewer than %llu bytes", path, (unsigned long long)bytes);
	}
	fseeko(f, 0, SEEK_SET);
	std::lock_guard<std::mutex> lk(ctx->mu);
	HJ_HIP(ctx, hipSetDevice(ctx->device));
	void *stage[2] = {nullptr, nullptr};
	hipEvent_t done[2] = {nullptr, nullptr};
	int rc = HJ_OK;
	for (int b = 0; b < 2 && rc == HJ_OK; b++) {
		if (hipHostMalloc(&stage[b], kStage, hipHostMallocDefault) != hipSuccess ||
		    hipEventCreateWithFlags(&done[b], hipEventDisableTiming) != hipSuccess)
			rc = hj_fail(ctx, HJ_ENOMEM, "pinned staging buffers");
	}
	// read chunk k into buffer k % 2 while chunk k - 1 is copying
	for (uint64_t off = 0, k = 0; rc == HJ_OK && off < bytes; off += kStage, k++) {
		const int b = (int)(k & 1);
		const size_t n = (size_t)std::min<uint64_t>(kStage, bytes - off);
		if (k >= 2 && hipEventSynchronize(done[b]) != hipSuccess) {
			rc = hj_fail(ctx, HJ_EHIP, "staging copy failed");
			break;
		}
		if (fread(stage[b], 1, n, f) != n) {
			rc = hj_fail(ctx, HJ_EINVAL, "short read from %s", path);
			break;
		}
		if (hipMemcpyAsync((char *)dev + off, stage[b], n, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
		    hipEventRecord(done[b], ctx->stream) != hipSuccess)
			rc = hj_fail(ctx, HJ_EHIP, "copy to device failed");
	}
	(void)hipStreamSynchronize(ctx->stream);
	for (int b = 0; b < 2; b++) {
		if (done[b]) (void)hipEventDestroy(done[b]);
		if (stage[b]) (void)hipHostFree(stage[b]);
	}
	fclose(f);
	return rc;
}

int hj_bin_write(hj_ctx_t *ctx, const char *path, const void *dev, uint64_t tuple_num, uint32_t tuple_size) {
	if (!ctx || !path || (tuple_num && !dev)) return HJ_EINVAL;
	const uint64_t bytes = tuple_num * tuple_size;
	FILE *f = fopen(path, "wb");
	if (!f) return hj_fail(ctx, HJ_EINVAL, "cannot create %s: %s", path, strerror(errno));
	std::lock_guard<std::mutex> lk(ctx->mu);
	HJ_HIP(ctx, hipSetDevice(ctx->device));
	void *stage = nullptr;
	int rc = HJ_OK;
	if (hipHostMalloc(&stage, kStage, hipHostMallocDefault) != hipSuccess)
		rc = hj_fail(ctx, HJ_ENOMEM, "pinned staging buffer");
	for (uint64_t off = 0; rc == HJ_OK && off < bytes; off += kStage) {
		const size_t n = (size_t)std::min<uint64_t>(kStage, bytes - off);
		if (hipMemcpyAsync(stage, (const char *)dev + off, n, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
		    hipStreamSynchronize(ctx->stream) != hipSuccess) {
			rc = hj_fail(ctx, HJ_EHIP, "copy from device failed");
			break;
		}
		if (fwrite(stage, 1, n, f) != n) rc = hj_fail(ctx, HJ_EINVAL, "short write to %s", path);
	}
	if (stage) (void)hipHostFree(stage);
	if (fclose(f) && rc == HJ_OK) rc = hj_fail(ctx, HJ_EINVAL, "closing %s failed", path);
	return rc;
}

}  // extern "C"
