// hj_gen.hip -- seeded relation generation on the device.
//
// Every tuple is a pure function of (spec, row index) through hjgen.h, so the
// device fills relations of any size in one pass at HBM speed and the host
// oracle reproduces the same bytes.  Replaces bin/gen (src/gen.c:29-158),
// which shuffles host arrays with glibc rand().

#include "hj_internal.h"
#include "hjgen.h"

#include <vector>

template <int WORDS>
__global__ __launch_bounds__(256) void k_gen_rel(hj_gen_state_t g, int side, const double *__restrict__ cdf,
                                                 uint64_t n, uint4 *__restrict__ out) {
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
		uint64_t w[WORDS];
		if (side == 0)
			hj_gen_r_tuple(&g, i, w);
		else
			hj_gen_s_tuple(&g, i, cdf, w);
#pragma unroll
		for (int k = 0; k < WORDS / 2; k++) {
			uint4 v;
			v.x = (uint32_t)w[2 * k];
			v.y = (uint32_t)(w[2 * k] >> 32);
			v.z = (uint32_t)w[2 * k + 1];
			v.w = (uint32_t)(w[2 * k + 1] >> 32);
			out[i * (WORDS / 2) + k] = v;
		}
	}
}

template <int WORDS>
static void launch_gen(hj_ctx_t *ctx, const hj_gen_state_t &g, int side, const double *cdf, uint64_t n,
                       void *out) {
	if (n == 0) return;
	uint64_t blocks = (n + 255) / 256;
	uint64_t cap = (uint64_t)ctx->num_cu * 32;
	if (blocks > cap) blocks = cap;
	hipLaunchKernelGGL(k_gen_rel<WORDS>, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, g, side, cdf, n,
	                   (uint4 *)out);
}

int hj_launch_generate(hj_ctx_t *ctx, const hj_gen_spec_t *spec, void *r, void *s) {
	hj_gen_state_t g = hj_gen_state(spec);
	double *cdf = nullptr;
	if (spec->workload == HJ_WL_ZIPF && spec->s_num) {
		std::vector<double> h(spec->r_num);
		hj_zipf_cdf_build(h.data(), spec->r_num, spec->theta);
		HJ_HIP(ctx, hipMalloc(&cdf, h.size() * sizeof(double)));
		hipError_t e = hipMemcpy(cdf, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice);
		if (e != hipSuccess) {
			(void)hipFree(cdf);
			return hj_fail(ctx, HJ_EHIP, "zipf table copy: %s", hipGetErrorString(e));
		}
	}
	for (int side = 0; side < 2; side++) {
		uint64_t n = side ? spec->s_num : spec->r_num;
		void *out = side ? s : r;
		switch (spec->width) {
		case 16: launch_gen<2>(ctx, g, side, cdf, n, out); break;
		case 32: launch_gen<4>(ctx, g, side, cdf, n, out); break;
		case 64: launch_gen<8>(ctx, g, side, cdf, n, out); break;
		case 128: launch_gen<16>(ctx, g, side, cdf, n, out); break;
		case 256: launch_gen<32>(ctx, g, side, cdf, n, out); break;
		default: launch_gen<64>(ctx, g, side, cdf, n, out); break;
		}
	}
	hipError_t e = hipGetLastError();
	if (cdf) {
		(void)hipStreamSynchronize(ctx->stream);
		(void)hipFree(cdf);
	}
	if (e != hipSuccess) return hj_fail(ctx, HJ_EHIP, "generator launch: %s", hipGetErrorString(e));
	return HJ_OK;
}
