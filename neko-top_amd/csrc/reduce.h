// reduce.h -- deterministic reductions for gfx950 (64-wide wavefronts).
//
// Every sum is taken with a fixed tree: a butterfly over the 64 lanes, then
// the wavefront partials of a block in order, then (in a one-block finalise
// kernel) the block partials in a fixed strided order.  For a given problem
// size the result is the same bit pattern on every run.
#pragma once
#include <hip/hip_runtime.h>

namespace nekoamd {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum over a block of BS threads (BS a multiple of 64).  sh: BS/64 doubles of
// LDS.  The result is valid on thread 0.
template <int BS>
__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < BS / 64; ++q) s += sh[q];
  }
  __syncthreads();
  return s;
}

// Sum of n partials by one block of BS threads; valid on thread 0.
template <int BS>
__device__ __forceinline__ double sum_partials(const double* __restrict__ p,
                                               int64_t n, double* sh) {
  double s = 0.0;
  for (int64_t q = threadIdx.x; q < n; q += BS) s += p[q];
  return block_sum<BS>(s, sh);
}

constexpr int kRedBlock = 256;   // threads per block of a reducing kernel
constexpr int kRedGrid = 2048;   // max blocks (= partials) of one reduction
constexpr int kFinBlock = 1024;  // threads of the finalise kernel

}  // namespace nekoamd
