// mdot.h -- streaming passes over several n-vectors at once: mult-weighted
// inner products <a, V_i>, i < m (mdots_k), and the point-pair helpers the
// linear-combination passes (projection, GMRES) share.
//
// <a, V_i>_mult for the m vectors V_i = V + i n (i < m <= 64) read a and
// mult once: every thread keeps one accumulator per vector, so the pass
// moves (m + 2) x 8 bytes per point instead of 3 m x 8 for m separate
// products.  Block partials go to part[i * G + blk] (G = the grid); a
// one-block-per-vector finalise sums them in a fixed order.
//
// When n is even and every base address is 16-byte aligned, a thread works
// on point pairs (2q, 2q + 1) with 16-byte loads; the basis vectors are
// streamed with nontemporal loads (each is read once per pass).  Measured on
// MI355X at n = 16.8 M, m = 15: 355 us (6.4 TB/s) against 587 us for one
// point per thread with ordinary loads on 2048 blocks; m = 20: 432 us
// against 687 us.  Loads go in fenced chunks of 8 vectors, so a thread has 8
// (not m) loads in flight and stays within a few dozen registers.
#pragma once
#include <cstdint>

#include "internal.h"
#include "reduce.h"

namespace nekoamd {

// Blocks of a multi-vector pass: 4096 x 256 threads (16 waves per CU).
constexpr int kMultiGrid = 4096;

// True when n-vectors based at each pointer can be read as 16-byte pairs.
static inline bool pairs_ok(int64_t n, const void* p0, const void* p1 = nullptr,
                            const void* p2 = nullptr, const void* p3 = nullptr,
                            const void* p4 = nullptr) {
  auto al = [](const void* p) {
    return p == nullptr || ((uintptr_t)p & 15) == 0;
  };
  return n % 2 == 0 && al(p0) && al(p1) && al(p2) && al(p3) && al(p4);
}

// The multiplicity weights mult[p] of a pass, stored as FP64 values
// (TM = double) or as the counts c = 1 / mult in one byte (TM = uint8_t,
// where every mult is exactly 1.0 / c for an integer c <= 255, which the
// caller has checked): 1.0 / c is then the same double, so the sums are
// the same bit for bit, for 1 byte per point instead of 8.
__device__ __forceinline__ double mult_at(const double* m, int64_t p) {
  return m[p];
}
__device__ __forceinline__ double mult_at(const uint8_t* m, int64_t p) {
  return 1.0 / (double)m[p];
}
__device__ __forceinline__ double2 mult_pair(const double* m, int64_t q) {
  return ((const double2*)m)[q];
}
__device__ __forceinline__ double2 mult_pair(const uint8_t* m, int64_t q) {
  const uchar2 c = ((const uchar2*)m)[q];
  return make_double2(1.0 / (double)c.x, 1.0 / (double)c.y);
}

// Nontemporal load of the pair p[0], p[1] (p 16-byte aligned).
__device__ __forceinline__ double2 ld_nt2(const double* p) {
  const double2* q = (const double2*)p;
  double2 v;
  v.x = __builtin_nontemporal_load(&q->x);
  v.y = __builtin_nontemporal_load(&q->y);
  return v;
}

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
// vectors V_i = V + i n: one pass over a and mult for all of them.  P2:
// point pairs (n even, 16-byte aligned bases).
template <int KM, bool P2, typename TM = double>
__global__ void __launch_bounds__(256)
    mdots_k(int64_t n, const double* __restrict__ a,
            const double* __restrict__ V, int m,
            const TM* __restrict__ mult, double* __restrict__ part) {
  __shared__ double sh[KM][4];
  double s[KM];
#pragma unroll
  for (int i = 0; i < KM; ++i) s[i] = 0.0;
  if constexpr (P2) {
    const int64_t n2 = n / 2;
    const double2* a2 = (const double2*)a;
    for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < n2;
         q += (int64_t)gridDim.x * 256) {
      const double2 av = a2[q], mv = mult_pair(mult, q);
      const double w0 = mv.x * av.x, w1 = mv.y * av.y;
#pragma unroll
      for (int c = 0; c < KM; c += 8) {
        if (c < m) {
#pragma unroll
          for (int i = c; i < c + 8; ++i)
            if (i < m) {
              const double2 v = ld_nt2(V + (int64_t)i * n + 2 * q);
              s[i] += w0 * v.x + w1 * v.y;
            }
          asm volatile("" ::: "memory");
        }
      }
    }
  } else {
    for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < n;
         p += (int64_t)gridDim.x * 256) {
      const double w = mult_at(mult, p) * a[p];
#pragma unroll
      for (int c = 0; c < KM; c += 8) {
        if (c < m) {
#pragma unroll
          for (int i = c; i < c + 8; ++i)
            if (i < m)
              s[i] += w * __builtin_nontemporal_load(V + (int64_t)i * n + p);
          asm volatile("" ::: "memory");
        }
      }
    }
  }
  block_partials<KM>(s, m, sh, part);
}

template <int KM, typename TM>
static inline void mdots_launch_km(int64_t n, const double* a, const double* V,
                                   int m, const TM* mult, double* part,
                                   int G, bool p2) {
  const dim3 g(G), b(256);
  if (p2)
    hipLaunchKernelGGL((mdots_k<KM, true, TM>), g, b, 0, stream(), n, a, V, m,
                       mult, part);
  else
    hipLaunchKernelGGL((mdots_k<KM, false, TM>), g, b, 0, stream(), n, a, V, m,
                       mult, part);
}

// Launch mdots_k on the library stream with G blocks of 256 threads (mult as
// FP64 values or one-byte counts, mult_at).
template <typename TM>
static inline void mdots_launch(int64_t n, const double* a, const double* V,
                                int m, const TM* mult, double* part,
                                int G) {
  // pairs: 16-byte aligned FP64 vectors; the counts are read as byte pairs
  const bool p2 = pairs_ok(n, a, V, sizeof(TM) == 8 ? (const void*)mult : nullptr) &&
                  ((uintptr_t)mult & (2 * sizeof(TM) - 1)) == 0;
  if (m <= 8)
    mdots_launch_km<8>(n, a, V, m, mult, part, G, p2);
  else if (m <= 16)
    mdots_launch_km<16>(n, a, V, m, mult, part, G, p2);
  else if (m <= 24)
    mdots_launch_km<24>(n, a, V, m, mult, part, G, p2);
  else if (m <= 32)
    mdots_launch_km<32>(n, a, V, m, mult, part, G, p2);
  else
    mdots_launch_km<64>(n, a, V, m, mult, part, G, p2);
}

// s += sum_{i<k} c[i] V_i at the point pair (2q, 2q + 1) (P2) or the point
// q (in s.x), for the vectors V_i = V + i n; c in LDS.  Chunks of 4 vectors,
// fenced.  TWO: the same coefficients also combine W_i into t (a basis and
// its images under the operator in one pass).  Measured at n = 16.8 M, two
// families of 15: 764 us (5.8 TB/s) against 864 us for one point per thread
// with an unchunked loop.
template <int KM, bool P2, bool TWO>
__device__ __forceinline__ void comb_point(const double* c,
                                           const double* __restrict__ V,
                                           const double* __restrict__ W,
                                           int64_t n, int k, int64_t q,
                                           double2& s, double2& t) {
#pragma unroll
  for (int c0 = 0; c0 < KM; c0 += 4) {
    if (c0 < k) {
#pragma unroll
      for (int i = c0; i < c0 + 4; ++i)
        if (i < k) {
          const double ci = c[i];
          if constexpr (P2) {
            const double2 v = ld_nt2(V + (int64_t)i * n + 2 * q);
            s.x += ci * v.x;
            s.y += ci * v.y;
            if constexpr (TWO) {
              const double2 w = ld_nt2(W + (int64_t)i * n + 2 * q);
              t.x += ci * w.x;
              t.y += ci * w.y;
            }
          } else {
            s.x += ci * __builtin_nontemporal_load(V + (int64_t)i * n + q);
            if constexpr (TWO)
              t.x += ci * __builtin_nontemporal_load(W + (int64_t)i * n + q);
          }
        }
      asm volatile("" ::: "memory");
    }
  }
}

}  // namespace nekoamd
