// pdc_util.hip — device-wide scan, scalar read-back and kernel timing.
#include "pdc_internal.h"

namespace pdc {

// Exclusive scan of n ints by one 1024-thread workgroup; out[n] = total.
// With n_dev, n is min(n, *n_dev) (a count known on the device only).
__global__ __launch_bounds__(1024) void k_scan_exclusive(const int* __restrict__ in,
                                                         int* __restrict__ out, int n,
                                                         const int* __restrict__ n_dev) {
	__shared__ int wsum[16];
	__shared__ int carry;
	if (n_dev) n = min(n, *n_dev);
	const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
	if (tid == 0) carry = 0;
	__syncthreads();
	for (int base = 0; base < n; base += 1024) {
		const int i = base + tid;
		const int v = i < n ? in[i] : 0;
		int x = v;
#pragma unroll
		for (int off = 1; off < 64; off <<= 1) {
			const int y = __shfl_up(x, off, 64);
			if (lane >= off) x += y;
		}
		if (lane == 63) wsum[wid] = x;
		__syncthreads();
		if (wid == 0) {
			int s = lane < 16 ? wsum[lane] : 0;
#pragma unroll
			for (int off = 1; off < 16; off <<= 1) {
				const int y = __shfl_up(s, off, 64);
				if (lane >= off) s += y;
			}
			if (lane < 16) wsum[lane] = s;
		}
		__syncthreads();
		const int before = wid > 0 ? wsum[wid - 1] : 0;
		if (i < n) out[i] = carry + before + x - v;
		__syncthreads();
		if (tid == 0) carry += wsum[15];
		__syncthreads();
	}
	if (tid == 0) out[n] = carry;
}

void scan_exclusive(pdc_context* ctx, const int* in, int* out, int n) {
	hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, ctx->stream, in, out, n, (const int*)nullptr);
	PDC_CHECK_LAUNCH();
}

void scan_exclusive_dev(pdc_context* ctx, const int* in, int* out, int cap, const int* n_dev) {
	hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, ctx->stream, in, out, cap, n_dev);
	PDC_CHECK_LAUNCH();
}

int read_int(pdc_context* ctx, const int* dptr) {
	int v = 0;
	PDC_HIP(hipMemcpyAsync(&v, dptr, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
	PDC_HIP(hipStreamSynchronize(ctx->stream));
	return v;
}

void sync_scalars(pdc_context* ctx) {
	PDC_HIP(hipMemcpyAsync(ctx->h_sc, ctx->d_sc.p, sizeof(Scalars), hipMemcpyDeviceToHost,
	                       ctx->stream));
	PDC_HIP(hipStreamSynchronize(ctx->stream));
}

void kt_begin(pdc_context* ctx, int which, hipEvent_t* ev) {
	*ev = nullptr;
	if (ctx->kt.which != which || which == PDC_KT_NONE) return;
	auto& kt = ctx->kt;
	if (kt.used + 2 > kt.pool.size()) {
		for (int k = 0; k < 512; ++k) {
			hipEvent_t e;
			PDC_HIP(hipEventCreate(&e));
			kt.pool.push_back(e);
		}
	}
	*ev = kt.pool[kt.used];
	PDC_HIP(hipEventRecord(*ev, ctx->stream));
}

void kt_end(pdc_context* ctx, int which, hipEvent_t ev0) {
	if (!ev0) return;
	auto& kt = ctx->kt;
	hipEvent_t e1 = kt.pool[kt.used + 1];
	PDC_HIP(hipEventRecord(e1, ctx->stream));
	kt.used += 2;
	kt.pair_launches.push_back(1);
	if (kt.used + 2 > 1024) kt_collect(ctx);
}

void kt_end_span(pdc_context* ctx, hipEvent_t ev0, int launches) {
	if (!ev0) return;
	auto& kt = ctx->kt;
	PDC_HIP(hipEventRecord(kt.pool[kt.used + 1], ctx->stream));
	kt.used += 2;
	kt.pair_launches.push_back(launches);
	if (kt.used + 2 > 1024) kt_collect(ctx);
}

void kt_end_nocollect(pdc_context* ctx, hipEvent_t ev0) {
	if (!ev0) return;
	auto& kt = ctx->kt;
	PDC_HIP(hipEventRecord(kt.pool[kt.used + 1], ctx->stream));
	kt.used += 2;
	kt.pair_launches.push_back(1);
}

void kt_collect(pdc_context* ctx) {
	auto& kt = ctx->kt;
	if (kt.used == 0) return;
	PDC_HIP(hipStreamSynchronize(ctx->stream));
	for (size_t k = 0; k + 1 < kt.used; k += 2) {
		float ms = 0;
		PDC_HIP(hipEventElapsedTime(&ms, kt.pool[k], kt.pool[k + 1]));
		kt.total_ms += ms;
		kt.launches += k / 2 < kt.pair_launches.size() ? kt.pair_launches[k / 2] : 1;
	}
	kt.used = 0;
	kt.pair_launches.clear();
}

}  // namespace pdc
