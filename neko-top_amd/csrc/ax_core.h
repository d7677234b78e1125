// ax_core.h -- the per-point body of the matrix-free Helmholtz operator,
// shared by the plain ax_helm kernels and the fused CG iteration.
//
// One workgroup per element, one thread per GLL point (t = i + LX*(j +
// LX*k)); at lx = 8 that is 512 threads, eight 64-wide wavefronts.  Each
// thread issues its own coalesced loads of u, G11..G33, h1 (h2, B) before
// the first barrier, so all of an element's loads are in flight together;
// the contractions run out of LDS tiles of u and of the three fluxes, and
// the derivative matrix sits in LDS too.  Few registers per thread (about
// 60) let eight waves per SIMD hide HBM latency.
//
//   w = sum_a D_a^T (h1 sum_b G_ab D_b u) [+ h2 B u]
//
// TA: arithmetic type (LDS tiles, contractions); the caller converts its
// storage types to TA when it fills AxPoint and su.
//
// The operator data (G, h1, h2, B) are pure streams, read once per call and
// never again before the next call evicts them, so they are loaded
// nontemporally (measured at 32^3 p=7: 203 us against 217 us per FP64
// call, bit-identical results).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace nekoamd {

// Operator data of one point, loaded before the first barrier.
template <typename TA>
struct AxPoint {
  TA g11, g12, g13, g22, g23, g33, h1, h2b;  // h2b = h2 * B (H2 only)
};

template <typename TA, typename TH, typename TG, bool H2>
__device__ __forceinline__ AxPoint<TA> ax_load(
    int64_t g, const TH* __restrict__ h1, const TH* __restrict__ h2,
    const TG* __restrict__ g11, const TG* __restrict__ g12,
    const TG* __restrict__ g13, const TG* __restrict__ g22,
    const TG* __restrict__ g23, const TG* __restrict__ g33,
    const TG* __restrict__ bm) {
  AxPoint<TA> a;
  a.g11 = (TA)__builtin_nontemporal_load(g11 + g);
  a.g12 = (TA)__builtin_nontemporal_load(g12 + g);
  a.g13 = (TA)__builtin_nontemporal_load(g13 + g);
  a.g22 = (TA)__builtin_nontemporal_load(g22 + g);
  a.g23 = (TA)__builtin_nontemporal_load(g23 + g);
  a.g33 = (TA)__builtin_nontemporal_load(g33 + g);
  a.h1 = (TA)__builtin_nontemporal_load(h1 + g);
  if constexpr (H2)
    a.h2b = (TA)__builtin_nontemporal_load(h2 + g) *
            (TA)__builtin_nontemporal_load(bm + g);
  else
    a.h2b = TA(0);
  return a;
}

// su holds the element's u (every thread stored its value) and sD the
// derivative matrix D[i*LX+l] = l_l'(z_i); the caller synchronised.  sr, ss,
// st: flux tiles.  Returns w at point t (u = the thread's own value).
template <int LX, typename TA, bool H2>
__device__ __forceinline__ TA ax_point(TA u, const AxPoint<TA>& a,
                                       const TA* su, TA* sr, TA* ss, TA* st,
                                       const TA* sD, int t) {
  constexpr int NS = LX * LX;
  const int i = t % LX, j = (t / LX) % LX, k = t / NS;
  TA ur = 0, us = 0, ut = 0;
#pragma unroll
  for (int l = 0; l < LX; ++l) {
    ur += sD[i * LX + l] * su[k * NS + j * LX + l];
    us += sD[j * LX + l] * su[k * NS + l * LX + i];
    ut += sD[k * LX + l] * su[l * NS + j * LX + i];
  }
  sr[t] = a.h1 * (a.g11 * ur + a.g12 * us + a.g13 * ut);
  ss[t] = a.h1 * (a.g12 * ur + a.g22 * us + a.g23 * ut);
  st[t] = a.h1 * (a.g13 * ur + a.g23 * us + a.g33 * ut);
  __syncthreads();
  TA s = 0;
#pragma unroll
  for (int l = 0; l < LX; ++l) {
    s += sD[l * LX + i] * sr[k * NS + j * LX + l];
    s += sD[l * LX + j] * ss[k * NS + l * LX + i];
    s += sD[l * LX + k] * st[l * NS + j * LX + i];
  }
  if constexpr (H2) s += a.h2b * u;
  return s;
}

// ax_point at LX * LX = 64, where a wavefront is one k-plane of the element:
// the t-direction's row and column of D (D[k][l], D[l][k]) are wave-uniform
// and come from global memory by scalar loads instead of from LDS (80 LDS
// reads per point instead of 96); dxg holds D in TA.  The same products
// summed in the same order, so the result is ax_point's bit for bit.
template <int LX, typename TA, bool H2>
__device__ __forceinline__ TA ax_point_sk(TA u, const AxPoint<TA>& a,
                                          const TA* su, TA* sr, TA* ss, TA* st,
                                          const TA* sD,
                                          const TA* __restrict__ dxg, int t) {
  constexpr int NS = LX * LX;
  static_assert(NS == 64, "a wavefront is one k-plane");
  const int i = t % LX, j = (t / LX) % LX;
  const int k = __builtin_amdgcn_readfirstlane(t / NS);
  TA ur = 0, us = 0, ut = 0;
#pragma unroll
  for (int l = 0; l < LX; ++l) {
    ur += sD[i * LX + l] * su[k * NS + j * LX + l];
    us += sD[j * LX + l] * su[k * NS + l * LX + i];
    ut += dxg[k * LX + l] * su[l * NS + j * LX + i];
  }
  sr[t] = a.h1 * (a.g11 * ur + a.g12 * us + a.g13 * ut);
  ss[t] = a.h1 * (a.g12 * ur + a.g22 * us + a.g23 * ut);
  st[t] = a.h1 * (a.g13 * ur + a.g23 * us + a.g33 * ut);
  __syncthreads();
  TA s = 0;
#pragma unroll
  for (int l = 0; l < LX; ++l) {
    s += sD[l * LX + i] * sr[k * NS + j * LX + l];
    s += sD[l * LX + j] * ss[k * NS + l * LX + i];
    s += dxg[l * LX + k] * st[l * NS + j * LX + i];
  }
  if constexpr (H2) s += a.h2b * u;
  return s;
}

// Sum of v over the workgroup of NP = LX^3 threads; valid on thread 0.
// sd: NP doubles of LDS.  A fixed order: wavefront butterflies, then the
// wavefront sums in order when NP is a multiple of 64; else a serial sum.
template <int NP>
__device__ __forceinline__ double element_sum(double v, double* sd) {
  if constexpr (NP % 64 == 0) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) sd[threadIdx.x >> 6] = v;
    __syncthreads();
    double a = 0.0;
    if (threadIdx.x == 0)
      for (int q = 0; q < NP / 64; ++q) a += sd[q];
    return a;
  } else {
    sd[threadIdx.x] = v;
    __syncthreads();
    double a = 0.0;
    if (threadIdx.x == 0)
      for (int q = 0; q < NP; ++q) a += sd[q];
    return a;
  }
}

}  // namespace nekoamd
