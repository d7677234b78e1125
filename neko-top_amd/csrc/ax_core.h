// ax_core.h -- the per-element body of the matrix-free Helmholtz operator,
// shared by the plain ax_helm kernels and the fused CG iteration.
//
// One workgroup per element, one thread per (i,j) column (t = i + LX*j)
// walking k.  On entry the thread's column of u is in uc[] and the whole
// element is in the LDS tile su (the caller stored it and synchronised).
// On exit wk[k] holds w at (i,j,k):
//   w = sum_a D_a^T (h1 sum_b G_ab D_b u) [+ h2 B u]
// sr, ss: LDS tiles for the r- and s-direction fluxes.
// TA: arithmetic type (of uc, the tiles and wk); TH: storage type of h1, h2;
// TG: storage type of the geometric factors and B.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace nekoamd {

template <int LX, typename T, typename TH, typename TG, bool H2>
__device__ __forceinline__ void ax_element(
    const T (&uc)[LX], const T* su, T* sr, T* ss, int t, int64_t eb,
    const double* __restrict__ dx, const TH* __restrict__ h1,
    const TH* __restrict__ h2, const TG* __restrict__ g11,
    const TG* __restrict__ g12, const TG* __restrict__ g13,
    const TG* __restrict__ g22, const TG* __restrict__ g23,
    const TG* __restrict__ g33, const TG* __restrict__ bm, T (&wk)[LX]) {
  constexpr int NS = LX * LX;
  const int i = t % LX, j = t / LX;
  T Di[LX], Dj[LX], DiT[LX], DjT[LX];
#pragma unroll
  for (int l = 0; l < LX; ++l) {
    Di[l] = (T)dx[i * LX + l];
    Dj[l] = (T)dx[j * LX + l];
    DiT[l] = (T)dx[l * LX + i];
    DjT[l] = (T)dx[l * LX + j];
  }
  T wt[LX];
#pragma unroll
  for (int k = 0; k < LX; ++k) {
    const int64_t g = eb + k * NS;
    const T G11 = (T)g11[g], G12 = (T)g12[g], G13 = (T)g13[g];
    const T G22 = (T)g22[g], G23 = (T)g23[g], G33 = (T)g33[g];
    const T hh = (T)h1[g];
    T ur = 0, us = 0, ut = 0;
#pragma unroll
    for (int l = 0; l < LX; ++l) {
      ur += Di[l] * su[k * NS + j * LX + l];
      us += Dj[l] * su[k * NS + l * LX + i];
      ut += (T)dx[k * LX + l] * uc[l];
    }
    sr[k * NS + t] = hh * (G11 * ur + G12 * us + G13 * ut);
    ss[k * NS + t] = hh * (G12 * ur + G22 * us + G23 * ut);
    wt[k] = hh * (G13 * ur + G23 * us + G33 * ut);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < LX; ++k) {
    T s = 0;
#pragma unroll
    for (int l = 0; l < LX; ++l) {
      s += DiT[l] * sr[k * NS + j * LX + l];
      s += DjT[l] * ss[k * NS + l * LX + i];
      s += (T)dx[l * LX + k] * wt[l];
    }
    if constexpr (H2) {
      const int64_t g = eb + k * NS;
      s += (T)h2[g] * (T)bm[g] * uc[k];
    }
    wk[k] = s;
  }
}

// Sum of v over the workgroup of NS = LX*LX threads; valid on thread 0.
// sd: NS doubles of LDS (used when NS != 64).
template <int NS>
__device__ __forceinline__ double element_sum(double v, double* sd) {
  if constexpr (NS == 64) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
  } else {
    sd[threadIdx.x] = v;
    __syncthreads();
    double a = 0.0;
    if (threadIdx.x == 0)
      for (int q = 0; q < NS; ++q) a += sd[q];
    return a;
  }
}

}  // namespace nekoamd
