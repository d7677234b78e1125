// coef.hip -- geometric factors of the SEM discretisation, built on the
// device from the element corner coordinates (what Neko's coef_t%init
// computes on the host before the reference's solves; consumed by every
// ax_t%compute, e.g. sources/adjoint/adjoint_pnpn.f90:716-723).
//
// One workgroup per element, one thread per (i,j) column walking k.
// Coordinates come from the trilinear map of the 8 corners, derivatives
// from the GLL derivative matrix, then
//   c_r = x_s x x_t, c_s = x_t x x_r, c_t = x_r x x_s, J = x_r . c_r,
//   G_ab = w_i w_j w_k (c_a . c_b) / J,  B = w_i w_j w_k J.
#include "internal.h"
#include "nekoamd.h"

namespace {

template <int LX>
__global__ void __launch_bounds__(LX * LX)
    coef_geo_k(double* __restrict__ xo, double* __restrict__ yo,
               double* __restrict__ zo, double* __restrict__ g11,
               double* __restrict__ g12, double* __restrict__ g13,
               double* __restrict__ g22, double* __restrict__ g23,
               double* __restrict__ g33, double* __restrict__ bm,
               double* __restrict__ jac, const double* __restrict__ corners,
               const double* __restrict__ zg, const double* __restrict__ wg,
               const double* __restrict__ dx) {
  constexpr int NS = LX * LX, NP = NS * LX;
  __shared__ double sx[3][NP];
  __shared__ double sc[24];
  const int e = blockIdx.x;
  const int t = threadIdx.x, i = t % LX, j = t / LX;
  for (int q = t; q < 24; q += NS) sc[q] = corners[(int64_t)e * 24 + q];
  __syncthreads();
  const int64_t base = (int64_t)e * NP + t;
  const double a0 = 0.5 * (1.0 - zg[i]), a1 = 0.5 * (1.0 + zg[i]);
  const double b0 = 0.5 * (1.0 - zg[j]), b1 = 0.5 * (1.0 + zg[j]);
#pragma unroll
  for (int k = 0; k < LX; ++k) {
    const double c0 = 0.5 * (1.0 - zg[k]), c1 = 0.5 * (1.0 + zg[k]);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double e00 = a0 * sc[0 * 3 + a] + a1 * sc[1 * 3 + a];
      const double e10 = a0 * sc[2 * 3 + a] + a1 * sc[3 * 3 + a];
      const double e01 = a0 * sc[4 * 3 + a] + a1 * sc[5 * 3 + a];
      const double e11 = a0 * sc[6 * 3 + a] + a1 * sc[7 * 3 + a];
      sx[a][k * NS + t] = c0 * (b0 * e00 + b1 * e10) + c1 * (b0 * e01 + b1 * e11);
    }
    if (xo) xo[base + k * NS] = sx[0][k * NS + t];
    if (yo) yo[base + k * NS] = sx[1][k * NS + t];
    if (zo) zo[base + k * NS] = sx[2][k * NS + t];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < LX; ++k) {
    double xr[3] = {0, 0, 0}, xs[3] = {0, 0, 0}, xt[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int m = 0; m < LX; ++m) {
        xr[a] += dx[i * LX + m] * sx[a][k * NS + j * LX + m];
        xs[a] += dx[j * LX + m] * sx[a][k * NS + m * LX + i];
        xt[a] += dx[k * LX + m] * sx[a][m * NS + t];
      }
    const double cr0 = xs[1] * xt[2] - xs[2] * xt[1];
    const double cr1 = xs[2] * xt[0] - xs[0] * xt[2];
    const double cr2 = xs[0] * xt[1] - xs[1] * xt[0];
    const double cs0 = xt[1] * xr[2] - xt[2] * xr[1];
    const double cs1 = xt[2] * xr[0] - xt[0] * xr[2];
    const double cs2 = xt[0] * xr[1] - xt[1] * xr[0];
    const double ct0 = xr[1] * xs[2] - xr[2] * xs[1];
    const double ct1 = xr[2] * xs[0] - xr[0] * xs[2];
    const double ct2 = xr[0] * xs[1] - xr[1] * xs[0];
    const double J = xr[0] * cr0 + xr[1] * cr1 + xr[2] * cr2;
    const double w3 = wg[i] * wg[j] * wg[k];
    const double s = w3 / J;
    const int64_t g = base + k * NS;
    g11[g] = s * (cr0 * cr0 + cr1 * cr1 + cr2 * cr2);
    g12[g] = s * (cr0 * cs0 + cr1 * cs1 + cr2 * cs2);
    g13[g] = s * (cr0 * ct0 + cr1 * ct1 + cr2 * ct2);
    g22[g] = s * (cs0 * cs0 + cs1 * cs1 + cs2 * cs2);
    g23[g] = s * (cs0 * ct0 + cs1 * ct1 + cs2 * ct2);
    g33[g] = s * (ct0 * ct0 + ct1 * ct1 + ct2 * ct2);
    bm[g] = w3 * J;
    if (jac) jac[g] = J;
  }
}

template <int LX>
int launch_coef(double* x, double* y, double* z, double* g11, double* g12,
                double* g13, double* g22, double* g23, double* g33, double* B,
                double* jac, const double* corners, const double* zg,
                const double* wg, const double* dx, int nelv) {
  hipLaunchKernelGGL(coef_geo_k<LX>, dim3(nelv), dim3(LX * LX), 0,
                     nekoamd::stream(), x, y, z, g11, g12, g13, g22, g23, g33,
                     B, jac, corners, zg, wg, dx);
  NK_LAUNCHED();
  return nekoamd::kOk;
}

}  // namespace

extern "C" int nekoamd_coef_geo(void* x_d, void* y_d, void* z_d, void* g11,
                                void* g12, void* g13, void* g22, void* g23,
                                void* g33, void* B, void* jac,
                                const void* corners_d, const void* zgll_d,
                                const void* wgll_d, const void* dx_d,
                                const int* lx, const int* nelv) {
  NK_INIT();
  NK_REQUIRE(lx && nelv && g11 && g12 && g13 && g22 && g23 && g33 && B &&
                 corners_d && zgll_d && wgll_d && dx_d,
             "null argument");
  NK_REQUIRE(*nelv >= 0, "negative nelv");
  if (*nelv == 0) return nekoamd::kOk;
  auto* c = (const double*)corners_d;
  auto* zg = (const double*)zgll_d;
  auto* wg = (const double*)wgll_d;
  auto* d = (const double*)dx_d;
#define NK_COEF_CASE(L)                                                     \
  case L:                                                                   \
    return launch_coef<L>((double*)x_d, (double*)y_d, (double*)z_d,         \
                          (double*)g11, (double*)g12, (double*)g13,         \
                          (double*)g22, (double*)g23, (double*)g33,         \
                          (double*)B, (double*)jac, c, zg, wg, d, *nelv);
  switch (*lx) {
    NK_COEF_CASE(2)
    NK_COEF_CASE(3)
    NK_COEF_CASE(4)
    NK_COEF_CASE(5)
    NK_COEF_CASE(6)
    NK_COEF_CASE(7)
    NK_COEF_CASE(8)
    NK_COEF_CASE(9)
    NK_COEF_CASE(10)
    default:
      return nekoamd::fail_arg(__func__, "lx must be in [2, 10]");
  }
#undef NK_COEF_CASE
}
