// mdot.h -- several mult-weighted inner products in one pass.
//
// <a, V_i>_mult for the m vectors V_i = V + i n (i < m <= 64) read a and
// mult once: every thread keeps one accumulator per vector, so the pass
// moves (m + 2) x 8 bytes per point instead of 3 m x 8 for m separate
// products.  Block partials go to part[i * G + blk] (G = the grid); a
// one-block-per-vector finalise sums them in a fixed order.
#pragma once
#include "internal.h"
#include "reduce.h"

namespace nekoamd {

// Block partials of KM sums (the first m meaningful) to part[i * G + blk]:
// a butterfly over each wavefront, then the 4 wavefronts in order.
template <int KM>
__device__ __forceinline__ void block_partials(const double (&s)[KM], int m,
                                               double (*sh)[4],
                                               double* __restrict__ part) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < KM; ++i)
    if (i < m) {
      const double v = wave_sum(s[i]);
      if (lane == 0) sh[i][wv] = v;
    }
  __syncthreads();
  if ((int)threadIdx.x < m) {
    const int i = threadIdx.x;
    part[(int64_t)i * gridDim.x + blockIdx.x] =
        ((sh[i][0] + sh[i][1]) + sh[i][2]) + sh[i][3];
  }
}

// part[i * G + blk] = block blk's partial of <a, V_i>_mult, i < m, for the m
// vectors V_i = V + i n: one pass over a and mult for all of them.
template <int KM>
__global__ void __launch_bounds__(256)
    mdots_k(int64_t n, const double* __restrict__ a,
                const double* __restrict__ V, int m,
                const double* __restrict__ mult, double* __restrict__ part) {
  __shared__ double sh[KM][4];
  double s[KM];
#pragma unroll
  for (int i = 0; i < KM; ++i) s[i] = 0.0;
  for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * 256) {
    const double w = mult[p] * a[p];
    // chunks of 8 vectors, fenced: 8 loads in flight per thread, not m
#pragma unroll
    for (int c = 0; c < KM; c += 8) {
      if (c < m) {
#pragma unroll
        for (int i = c; i < c + 8; ++i)
          if (i < m) s[i] += w * V[(int64_t)i * n + p];
        asm volatile("" ::: "memory");
      }
    }
  }
  block_partials<KM>(s, m, sh, part);
}

// Launch mdots_k on the library stream with G blocks of 256 threads.
static inline void mdots_launch(int64_t n, const double* a, const double* V,
                                int m, const double* mult, double* part,
                                int G) {
  const dim3 g(G), b(256);
  if (m <= 8)
    hipLaunchKernelGGL(mdots_k<8>, g, b, 0, stream(), n, a, V, m, mult, part);
  else if (m <= 16)
    hipLaunchKernelGGL(mdots_k<16>, g, b, 0, stream(), n, a, V, m, mult, part);
  else if (m <= 24)
    hipLaunchKernelGGL(mdots_k<24>, g, b, 0, stream(), n, a, V, m, mult, part);
  else if (m <= 32)
    hipLaunchKernelGGL(mdots_k<32>, g, b, 0, stream(), n, a, V, m, mult, part);
  else
    hipLaunchKernelGGL(mdots_k<64>, g, b, 0, stream(), n, a, V, m, mult, part);
}

}  // namespace nekoamd
