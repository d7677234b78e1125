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
// Mapping onto CDNA4: one workgroup per element, one thread per GLL point
// (512 threads = eight wavefronts at lx = 8); see ax_core.h.  Every field
// access is one coalesced wavefront load or store (the element is
// contiguous); u and the three fluxes are staged in LDS.  HBM traffic per
// call is the algorithmic minimum: each of u, G11..G33, h1 (h2, B) is read
// once and w written once.
#include "ax_core.h"
#include "internal.h"
#include "nekoamd.h"

namespace nekoamd {

template <int LX, typename T, typename TG, bool H2>
__global__ void __launch_bounds__(LX * LX * LX)
    ax_helm_k(T* __restrict__ w, const T* __restrict__ u,
              const double* __restrict__ dx, const T* __restrict__ h1,
              const T* __restrict__ h2, const TG* __restrict__ g11,
              const TG* __restrict__ g12, const TG* __restrict__ g13,
              const TG* __restrict__ g22, const TG* __restrict__ g23,
              const TG* __restrict__ g33, const TG* __restrict__ bm,
              const int* __restrict__ skip) {
  constexpr int NS = LX * LX, NP = NS * LX;
  __shared__ T su[NP];
  __shared__ T sr[NP];
  __shared__ T ss[NP];
  __shared__ T st[NP];
  __shared__ T sD[NS];
  if (skip != nullptr && *skip) return;
  const int t = threadIdx.x;
  const int64_t g = (int64_t)blockIdx.x * NP + t;
  if (t < NS) sD[t] = (T)dx[t];
  const T uu = __builtin_nontemporal_load(u + g);
  const AxPoint<T> a =
      ax_load<T, T, TG, H2>(g, h1, h2, g11, g12, g13, g22, g23, g33, bm);
  su[t] = uu;
  __syncthreads();
  __builtin_nontemporal_store(ax_point<LX, T, H2>(uu, a, su, sr, ss, st, sD, t),
                              w + g);
}

// Three fields through the same operator in one launch (Neko's
// ax_t%compute_vector): the operator data (G, h1, h2, B) are read once per
// element and applied to each field in turn, with exactly ax_helm_k's
// arithmetic per field, so each result equals the single-field call bit for
// bit.  A field left out by the host (active 0) or whose skip flag is set is
// left alone.
struct Ax3Vec {
  double* w[3];
  const double* u[3];
  const int* skip[3];
  int active[3];
  // with F: w = (F - g) - A u (the velocity residual, Neko's pnpn_vel_res
  // after its operator) instead of A u
  const double* F[3];
  const double* g[3];
};

template <int LX, bool H2>
__global__ void __launch_bounds__(LX * LX * LX)
    ax_helm3_k(Ax3Vec v, const double* __restrict__ dx,
               const double* __restrict__ h1, const double* __restrict__ h2,
               const double* __restrict__ g11, const double* __restrict__ g12,
               const double* __restrict__ g13, const double* __restrict__ g22,
               const double* __restrict__ g23, const double* __restrict__ g33,
               const double* __restrict__ bm) {
  constexpr int NS = LX * LX, NP = NS * LX;
  __shared__ double su[NP];
  __shared__ double sr[NP];
  __shared__ double ss[NP];
  __shared__ double st[NP];
  __shared__ double sD[NS];
  const int t = threadIdx.x;
  const int64_t g = (int64_t)blockIdx.x * NP + t;
  if (t < NS) sD[t] = dx[t];
  const AxPoint<double> a = ax_load<double, double, double, H2>(
      g, h1, h2, g11, g12, g13, g22, g23, g33, bm);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (!v.active[c] || (v.skip[c] != nullptr && *v.skip[c])) continue;
    const double uu = __builtin_nontemporal_load(v.u[c] + g);
    su[t] = uu;
    __syncthreads();
    double au;
    if constexpr (LX * LX == 64)
      au = ax_point_sk<LX, double, H2>(uu, a, su, sr, ss, st, sD, dx, t);
    else
      au = ax_point<LX, double, H2>(uu, a, su, sr, ss, st, sD, t);
    __builtin_nontemporal_store(
        v.F[c] != nullptr ? v.F[c][g] - v.g[c][g] - au : au, v.w[c] + g);
    __syncthreads();  // this field's flux reads before the next field's writes
  }
}

int ax_helm3_launch(double* const* w, const double* const* u,
                    const int* const* skip, const bool* active,
                    const double* dx, const double* h1, const double* h2,
                    const double* g11, const double* g12, const double* g13,
                    const double* g22, const double* g23, const double* g33,
                    const double* B, int lx, int nelv, bool ifh2,
                    const double* const* F, const double* const* G) {
  if (nelv == 0) return kOk;
  Ax3Vec v{};
  for (int c = 0; c < 3; ++c) {
    v.w[c] = w[c];
    v.u[c] = u[c];
    v.skip[c] = skip ? skip[c] : nullptr;
    v.active[c] = active == nullptr || active[c] ? 1 : 0;
    v.F[c] = F ? F[c] : nullptr;
    v.g[c] = F ? G[c] : nullptr;
  }
#define NK_AX3_CASE(L)                                                       \
  case L:                                                                    \
    if (ifh2)                                                                \
      nk_launch(ax_helm3_k<L, true>, dim3(nelv), dim3(L * L * L), 0, v, dx, \
                h1, h2, g11, g12, g13, g22, g23, g33, B);                    \
    else                                                                     \
      nk_launch(ax_helm3_k<L, false>, dim3(nelv), dim3(L * L * L), 0, v,    \
                dx, h1, h2, g11, g12, g13, g22, g23, g33, B);                \
    break;
  switch (lx) {
    NK_AX3_CASE(2)
    NK_AX3_CASE(3)
    NK_AX3_CASE(4)
    NK_AX3_CASE(5)
    NK_AX3_CASE(6)
    NK_AX3_CASE(7)
    NK_AX3_CASE(8)
    NK_AX3_CASE(9)
    NK_AX3_CASE(10)
    default:
      return fail_arg("ax_helm_vector", "lx must be in [2, 10]");
  }
#undef NK_AX3_CASE
  NK_LAUNCHED();
  return kOk;
}

template <typename T, typename TG>
int ax_helm_launch(T* w, const T* u, const double* dx, const T* h1,
                   const T* h2, const TG* g11, const TG* g12, const TG* g13,
                   const TG* g22, const TG* g23, const TG* g33, const TG* B,
                   int lx, int nelv, bool ifh2, const int* skip) {
  if (nelv == 0) return kOk;
#define NK_AX_CASE(L)                                                        \
  case L:                                                                    \
    if (ifh2)                                                                \
      hipLaunchKernelGGL((ax_helm_k<L, T, TG, true>), dim3(nelv),       \
                         dim3(L * L * L), 0, stream(), w, u, dx, h1, h2, g11, \
                         g12, g13, g22, g23, g33, B, skip);            \
    else                                                                     \
      hipLaunchKernelGGL((ax_helm_k<L, T, TG, false>), dim3(nelv),      \
                         dim3(L * L * L), 0, stream(), w, u, dx, h1, h2, g11, \
                         g12, g13, g22, g23, g33, B, skip);            \
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

template int ax_helm_launch<double, double>(
    double*, const double*, const double*, const double*, const double*,
    const double*, const double*, const double*, const double*, const double*,
    const double*, const double*, int, int, bool, const int*);

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
  return ax_helm_launch<T, TG>(
      (T*)w, (const T*)u, (const double*)dx, (const T*)h1, (const T*)h2,
      (const TG*)g11, (const TG*)g12, (const TG*)g13, (const TG*)g22,
      (const TG*)g23, (const TG*)g33, (const TG*)B, *lx, *nelv, *ifh2 != 0,
      nullptr);
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

int nekoamd_ax_helm_vector_f64(void* w1, void* w2, void* w3, const void* u1,
                               const void* u2, const void* u3, const void* dx,
                               const void* h1, const void* h2,
                               const void* g11, const void* g12,
                               const void* g13, const void* g22,
                               const void* g23, const void* g33,
                               const void* B, const int* lx, const int* nelv,
                               const int* ifh2) {
  NK_INIT();
  NK_REQUIRE(w1 && w2 && w3 && u1 && u2 && u3 && dx && h1 && g11 && g12 &&
                 g13 && g22 && g23 && g33 && lx && nelv && ifh2,
             "null argument");
  NK_REQUIRE(!*ifh2 || (h2 && B), "ifh2 set but h2 or B is null");
  NK_REQUIRE(*nelv >= 0, "negative nelv");
  NK_REQUIRE(*lx >= 2 && *lx <= 10, "lx must be in [2, 10]");
  double* const w[3] = {(double*)w1, (double*)w2, (double*)w3};
  const double* const u[3] = {(const double*)u1, (const double*)u2,
                              (const double*)u3};
  return ax_helm3_launch(w, u, nullptr, nullptr, (const double*)dx,
                         (const double*)h1,
                         (const double*)h2, (const double*)g11,
                         (const double*)g12, (const double*)g13,
                         (const double*)g22, (const double*)g23,
                         (const double*)g33, (const double*)B, *lx, *nelv,
                         *ifh2 != 0, nullptr, nullptr);
}

int nekoamd_pnpn_vel_res_vector(void* r1, void* r2, void* r3, const void* u1,
                                const void* u2, const void* u3,
                                const void* F1, const void* F2,
                                const void* F3, const void* g1,
                                const void* g2, const void* g3,
                                const void* dx, const void* h1,
                                const void* h2, const void* g11,
                                const void* g12, const void* g13,
                                const void* g22, const void* g23,
                                const void* g33, const void* B, const int* lx,
                                const int* nelv, const int* ifh2) {
  NK_INIT();
  NK_REQUIRE(r1 && r2 && r3 && u1 && u2 && u3 && F1 && F2 && F3 && g1 && g2 &&
                 g3 && dx && h1 && g11 && g12 && g13 && g22 && g23 && g33 &&
                 lx && nelv && ifh2,
             "null argument");
  NK_REQUIRE(!*ifh2 || (h2 && B), "ifh2 set but h2 or B is null");
  NK_REQUIRE(*nelv >= 0, "negative nelv");
  NK_REQUIRE(*lx >= 2 && *lx <= 10, "lx must be in [2, 10]");
  double* const w[3] = {(double*)r1, (double*)r2, (double*)r3};
  const double* const u[3] = {(const double*)u1, (const double*)u2,
                              (const double*)u3};
  const double* const F[3] = {(const double*)F1, (const double*)F2,
                              (const double*)F3};
  const double* const G[3] = {(const double*)g1, (const double*)g2,
                              (const double*)g3};
  return ax_helm3_launch(w, u, nullptr, nullptr, (const double*)dx,
                         (const double*)h1, (const double*)h2,
                         (const double*)g11, (const double*)g12,
                         (const double*)g13, (const double*)g22,
                         (const double*)g23, (const double*)g33,
                         (const double*)B, *lx, *nelv, *ifh2 != 0, F, G);
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
