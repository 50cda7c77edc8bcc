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

// Sums K values per thread over the workgroup; thread 0 receives the totals
// in out[] (other threads get unspecified values).  blockDim.x must be a
// multiple of 64 and at most 1024.  `lds` holds K * 16 values.
template <int K, class T>
__device__ __forceinline__ void block_sum(T (&v)[K], T* lds, T (&out)[K]) {
	const int lane = threadIdx.x & 63;
	const int wid = threadIdx.x >> 6;
	const int nw = blockDim.x >> 6;
#pragma unroll
	for (int k = 0; k < K; ++k) {
		T s = wave_sum(v[k]);
		if (lane == 0) lds[k * 16 + wid] = s;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
#pragma unroll
		for (int k = 0; k < K; ++k) {
			T s = 0;
			for (int w = 0; w < nw; ++w) s += lds[k * 16 + w];
			out[k] = s;
		}
	}
	__syncthreads();
}

// Sums part[0..n) in a fixed order by the first wave; all threads of the
// workgroup get the result (through `slot`, an LDS scalar).
template <class T>
__device__ __forceinline__ T block_reduce_array(const T* part, int n, T* slot) {
	if (threadIdx.x < 64) {
		T s = 0;
		for (int i = threadIdx.x; i < n; i += 64) s += part[i];
		s = wave_sum(s);
		if (threadIdx.x == 0) *slot = s;
	}
	__syncthreads();
	T r = *slot;
	__syncthreads();
	return r;
}

}  // namespace pdc
