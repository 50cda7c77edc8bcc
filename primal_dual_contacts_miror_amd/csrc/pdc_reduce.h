// pdc_reduce.h — fixed-order workgroup reductions (64-wide wavefronts).
//
// Every reduction here has an order fixed by the launch shape, so results do
// not vary from run to run; no floating-point atomics are used anywhere.
#pragma once

#include <hip/hip_runtime.h>

namespace pdc {

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
	return v;
}

template <class T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) {
		T o = __shfl_xor(v, off, 64);
		v = o > v ? o : v;
	}
	return v;
}

// ---- DPP forms (every lane of the wavefront active) -------------------------
// A lane permutation inside rows of 16 lanes through the data-parallel
// primitives' operand modifier: a VALU move instead of an LDS round trip.
constexpr int kDppXor1 = 0xB1;        // quad_perm [1, 0, 3, 2]
constexpr int kDppXor2 = 0x4E;        // quad_perm [2, 3, 0, 1]
constexpr int kDppMirror = 0x140;     // row_mirror: lane j of 16 reads lane 15 - j
constexpr int kDppHalfMirror = 0x141; // row_half_mirror: lane j of 8 reads lane 7 - j

template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
	const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
	const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
	return __hiloint2double(hi, lo);
}

// Sum over each group of 8 consecutive lanes, the same bits in every lane of
// the group (each step adds two partial sums that both lanes of a pair hold).
__device__ __forceinline__ double group8_sum(double v) {
	v += dpp_move<kDppXor1>(v);
	v += dpp_move<kDppXor2>(v);
	v += dpp_move<kDppHalfMirror>(v);
	return v;
}

// Sum over the wavefront in a fixed order, the same bits in every lane: 16-lane
// sums by DPP, then the four row sums read from lanes 0, 16, 32, 48.
__device__ __forceinline__ double wave_sum_dpp(double v) {
	v = group8_sum(v);
	v += dpp_move<kDppMirror>(v);
	auto lane_value = [&](int l) {
		return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
		                        __builtin_amdgcn_readlane(__double2loint(v), l));
	};
	return (lane_value(0) + lane_value(16)) + (lane_value(32) + lane_value(48));
}

constexpr int kPartLoads = 16;              // partial values per lane and pass

// Lane share of a partial-sum array of W values (one per workgroup tile of
// a scene): a pass issues the loads of every 64-value group that
// holds values (a wavefront-uniform condition; lanes past W in the last group
// read the last value, which the sum then drops), all in flight together;
// one pass covers 1024 tiles.  Groups wholly past W are not read: every
// workgroup reads the same array, and repeated reads of its last line would
// all queue at one cache channel.  The wavefront sum of the returned values
// is the array total, in a fixed order.
__device__ __forceinline__ double partial_lane_sum(const double* __restrict__ src, int W) {
	const int lane = threadIdx.x & 63;
	double s = 0;
	for (int base = 0; base < W; base += 64 * kPartLoads) {
		double v[kPartLoads];
#pragma unroll
		for (int k = 0; k < kPartLoads; ++k) {
			v[k] = 0;
			if (base + 64 * k < W) v[k] = src[min(base + lane + 64 * k, W - 1)];
		}
#pragma unroll
		for (int k = 0; k < kPartLoads; ++k) s += base + lane + 64 * k < W ? v[k] : 0.0;
	}
	return s;
}

// The wavefront sum of a block sum: DPP for double (every lane active),
// shuffles otherwise.
template <class T>
__device__ __forceinline__ T wave_sum_full(T v) {
	return wave_sum(v);
}
template <>
__device__ __forceinline__ double wave_sum_full<double>(double v) {
	return wave_sum_dpp(v);
}

// Sums K values per thread over the workgroup; thread 0 receives the totals
// in out[] (other threads get unspecified values).  blockDim.x must be a
// multiple of 64 and at most 1024, and every lane of a calling wavefront
// active.  `lds` holds K * 16 values.
template <int K, class T>
__device__ __forceinline__ void block_sum(T (&v)[K], T* lds, T (&out)[K]) {
	const int lane = threadIdx.x & 63;
	const int wid = threadIdx.x >> 6;
	const int nw = blockDim.x >> 6;
#pragma unroll
	for (int k = 0; k < K; ++k) {
		T s = wave_sum_full(v[k]);
		if (lane == 0) lds[k * 16 + wid] = s;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		// the 16 reads of a value issued together (slots past nw read but
		// not added)
#pragma unroll 1
		for (int k = 0; k < K; ++k) {
			T part[16];
#pragma unroll
			for (int w = 0; w < 16; ++w) part[w] = lds[k * 16 + w];
			T s = 0;
#pragma unroll
			for (int w = 0; w < 16; ++w) s = w < nw ? s + part[w] : s;
			out[k] = s;
		}
	}
	__syncthreads();
}

// The same for NW wavefronts per workgroup (a compile-time count), with the
// DPP wavefront sums (every lane active) and without the closing barrier:
// `lds` (K * NW values) must not be written again by the workgroup.
template <int K, int NW>
__device__ __forceinline__ void block_sum_last(double (&v)[K], double* lds, double (&out)[K]) {
	const int lane = threadIdx.x & 63;
	const int wid = threadIdx.x >> 6;
	double s[K];
#pragma unroll
	for (int k = 0; k < K; ++k) s[k] = wave_sum_dpp(v[k]);
	if (lane == 0) {
#pragma unroll
		for (int k = 0; k < K; ++k) lds[k * NW + wid] = s[k];
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		double part[K][NW];
#pragma unroll
		for (int k = 0; k < K; ++k)
#pragma unroll
			for (int w = 0; w < NW; ++w) part[k][w] = lds[k * NW + w];
#pragma unroll
		for (int k = 0; k < K; ++k) {
			double t = 0;
#pragma unroll
			for (int w = 0; w < NW; ++w) t += part[k][w];
			out[k] = t;
		}
	}
}

}  // namespace pdc
