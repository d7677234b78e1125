// ax_helm.hip -- matrix-free Helmholtz operator on gfx950.
//
//   w_e = sum_a D_a^T ( h1 * sum_b G_ab * D_b u_e )  [+ h2 * B * u_e]
//
// Replaces Neko's ax_helm_t%compute(w, u, coef, msh, Xh), the operator the
// reference builds with ax_helm_factory at
// sources/adjoint/adjoint_pnpn.f90:207 (velocity), :217 (pressure) and
// sources/mapping_functions/PDE_filter_mapping.f90:177, and which every
// Krylov iteration of the pressure / velocity / filter solves applies.
//
// Mapping onto CDNA4.  One workgroup per element with one thread per (i,j)
// column; at lx = 8 that is exactly one 64-wide wavefront.  A thread walks
// the k direction in registers, so
//   - every k-plane of u, G11..G33, h1 and w is one 512 B coalesced
//     wavefront access (the element is 4 KiB contiguous per field);
//   - the t-direction contractions (D_3 u, D_3^T wt) never touch LDS;
//   - the r- and s-direction contractions read the u tile and the wr / ws
//     tiles from LDS (12 KiB per element in FP64).
// The rows of D a thread needs (row i, row j, column i, column j) are held
// in registers; the k-indexed entries are wavefront-uniform.
// HBM traffic per call is the algorithmic minimum: each of u, G11..G33, h1
// (h2, B) is read once and w written once.
//
// The CG variant (DOT) also returns per element the partial sum of w.u,
// which for a continuous u that vanishes on the Dirichlet mask equals the
// mult-weighted <QQ^T w, u> that CG needs, so no extra pass over w and p.
#include "internal.h"
#include "nekoamd.h"

namespace nekoamd {

template <int LX, typename T, typename TG, bool H2, bool DOT>
__global__ void __launch_bounds__(LX * LX)
    ax_helm_k(T* __restrict__ w, const T* __restrict__ u,
              const double* __restrict__ dx, const T* __restrict__ h1,
              const T* __restrict__ h2, const TG* __restrict__ g11,
              const TG* __restrict__ g12, const TG* __restrict__ g13,
              const TG* __restrict__ g22, const TG* __restrict__ g23,
              const TG* __restrict__ g33, const TG* __restrict__ bm,
              double* __restrict__ dotp, const int* __restrict__ skip) {
  constexpr int NS = LX * LX, NP = NS * LX;
  __shared__ T su[NP];
  __shared__ T sr[NP];
  __shared__ T ss[NP];
  if (skip != nullptr && *skip) return;
  const int e = blockIdx.x;
  const int t = threadIdx.x, i = t % LX, j = t / LX;
  const int64_t eb = (int64_t)e * NP + t;

  T Di[LX], Dj[LX], DiT[LX], DjT[LX];
#pragma unroll
  for (int l = 0; l < LX; ++l) {
    Di[l] = (T)dx[i * LX + l];
    Dj[l] = (T)dx[j * LX + l];
    DiT[l] = (T)dx[l * LX + i];
    DjT[l] = (T)dx[l * LX + j];
  }

  T uc[LX];
#pragma unroll
  for (int k = 0; k < LX; ++k) {
    uc[k] = u[eb + k * NS];
    su[k * NS + t] = uc[k];
  }
  __syncthreads();

  T wt[LX];
#pragma unroll
  for (int k = 0; k < LX; ++k) {
    const int64_t g = eb + k * NS;
    const T G11 = (T)g11[g], G12 = (T)g12[g], G13 = (T)g13[g];
    const T G22 = (T)g22[g], G23 = (T)g23[g], G33 = (T)g33[g];
    const T hh = h1[g];
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

  double dot = 0.0;
#pragma unroll
  for (int k = 0; k < LX; ++k) {
    T s = 0;
#pragma unroll
    for (int l = 0; l < LX; ++l) {
      s += DiT[l] * sr[k * NS + j * LX + l];
      s += DjT[l] * ss[k * NS + l * LX + i];
      s += (T)dx[l * LX + k] * wt[l];
    }
    const int64_t g = eb + k * NS;
    if constexpr (H2) s += h2[g] * (T)bm[g] * uc[k];
    w[g] = s;
    if constexpr (DOT) dot += (double)s * (double)uc[k];
  }

  if constexpr (DOT) {
    if constexpr (NS == 64) {
      // one full wavefront: butterfly over the 64 lanes
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) dot += __shfl_xor(dot, o, 64);
      if (t == 0) dotp[e] = dot;
    } else {
      __shared__ double sd[NS];
      sd[t] = dot;
      __syncthreads();
      if (t == 0) {
        double a = 0.0;
        for (int q = 0; q < NS; ++q) a += sd[q];
        dotp[e] = a;
      }
    }
  }
}

template <typename T, typename TG, bool DOT>
int ax_helm_launch(T* w, const T* u, const double* dx, const T* h1,
                   const T* h2, const TG* g11, const TG* g12, const TG* g13,
                   const TG* g22, const TG* g23, const TG* g33, const TG* B,
                   int lx, int nelv, bool ifh2, double* dotp,
                   const int* skip) {
  if (nelv == 0) return kOk;
#define NK_AX_CASE(L)                                                        \
  case L:                                                                    \
    if (ifh2)                                                                \
      hipLaunchKernelGGL((ax_helm_k<L, T, TG, true, DOT>), dim3(nelv),       \
                         dim3(L * L), 0, stream(), w, u, dx, h1, h2, g11,    \
                         g12, g13, g22, g23, g33, B, dotp, skip);            \
    else                                                                     \
      hipLaunchKernelGGL((ax_helm_k<L, T, TG, false, DOT>), dim3(nelv),      \
                         dim3(L * L), 0, stream(), w, u, dx, h1, h2, g11,    \
                         g12, g13, g22, g23, g33, B, dotp, skip);            \
    break;
  switch (lx) {
    NK_AX_CASE(2)
    NK_AX_CASE(3)
    NK_AX_CASE(4)
    NK_AX_CASE(5)
    NK_AX_CASE(6)
    NK_AX_CASE(7)
    NK_AX_CASE(8)
    NK_AX_CASE(9)
    NK_AX_CASE(10)
    default:
      return fail_arg("ax_helm", "lx must be in [2, 10]");
  }
#undef NK_AX_CASE
  NK_LAUNCHED();
  return kOk;
}

template int ax_helm_launch<double, double, true>(
    double*, const double*, const double*, const double*, const double*,
    const double*, const double*, const double*, const double*, const double*,
    const double*, const double*, int, int, bool, double*, const int*);
template int ax_helm_launch<float, double, true>(
    float*, const float*, const double*, const float*, const float*,
    const double*, const double*, const double*, const double*, const double*,
    const double*, const double*, int, int, bool, double*, const int*);

}  // namespace nekoamd

using namespace nekoamd;

namespace {

template <typename T, typename TG>
int ax_entry(const char* fn, void* w, const void* u, const void* dx,
             const void* h1, const void* h2, const void* g11, const void* g12,
             const void* g13, const void* g22, const void* g23,
             const void* g33, const void* B, const int* lx, const int* nelv,
             const int* ifh2) {
  if (!initialised()) {
    set_error(fn, "nekoamd_device_init was not called");
    return kErrState;
  }
  if (!(w && u && dx && h1 && g11 && g12 && g13 && g22 && g23 && g33 && lx &&
        nelv && ifh2))
    return fail_arg(fn, "null argument");
  if (*ifh2 && !(h2 && B)) return fail_arg(fn, "ifh2 set but h2 or B is null");
  if (*nelv < 0) return fail_arg(fn, "negative nelv");
  return ax_helm_launch<T, TG, false>(
      (T*)w, (const T*)u, (const double*)dx, (const T*)h1, (const T*)h2,
      (const TG*)g11, (const TG*)g12, (const TG*)g13, (const TG*)g22,
      (const TG*)g23, (const TG*)g33, (const TG*)B, *lx, *nelv, *ifh2 != 0,
      nullptr, nullptr);
}

}  // namespace

extern "C" {

int nekoamd_ax_helm_f64(void* w, const void* u, const void* dx, const void* h1,
                        const void* h2, const void* g11, const void* g12,
                        const void* g13, const void* g22, const void* g23,
                        const void* g33, const void* B, const int* lx,
                        const int* nelv, const int* ifh2) {
  return ax_entry<double, double>(__func__, w, u, dx, h1, h2, g11, g12, g13,
                                  g22, g23, g33, B, lx, nelv, ifh2);
}

int nekoamd_ax_helm_f32(void* w, const void* u, const void* dx, const void* h1,
                        const void* h2, const void* g11, const void* g12,
                        const void* g13, const void* g22, const void* g23,
                        const void* g33, const void* B, const int* lx,
                        const int* nelv, const int* ifh2) {
  return ax_entry<float, double>(__func__, w, u, dx, h1, h2, g11, g12, g13,
                                 g22, g23, g33, B, lx, nelv, ifh2);
}

int nekoamd_ax_helm_f32g32(void* w, const void* u, const void* dx,
                           const void* h1, const void* h2, const void* g11,
                           const void* g12, const void* g13, const void* g22,
                           const void* g23, const void* g33, const void* B,
                           const int* lx, const int* nelv, const int* ifh2) {
  return ax_entry<float, float>(__func__, w, u, dx, h1, h2, g11, g12, g13, g22,
                                g23, g33, B, lx, nelv, ifh2);
}

}  // extern "C"
