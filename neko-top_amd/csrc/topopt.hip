// topopt.hip -- the topology-optimisation kernels of neko-top on gfx950.
//
//   Brinkman source term  f -= chi u         simple_brinkman_source_term.f90:139-153
//   permeability force    -(k1 + (k0-k1)(q+1) rho/(q+rho))   design.f90:316-348
//   RAMP map + chain rule                     RAMP_mapping.f90:182-267
//   linear map + chain rule                   linear_mapping.f90:104-136
//   dF/dchi                                    minimum_dissipation_objective_function.f90:260-301
//   adjoint lube forcing  f_adj += K chi u     adjoint_lube_source_term.f90:173-206
//   dissipation objective (gradients fused)   minimum_dissipation_objective_function.f90:186-254
//   volume constraint and its sensitivity     volume_constraint.f90:177-242
//
// The elementwise terms are single fused streams (the reference chains 3 to
// 7 field operations for each: e.g. the sensitivity is col3, addcol3 x2,
// cmult, col3, addcol3 x2, add2s2 = 8 passes; here it is one read of six
// fields and one write).  The objective computes the three velocity
// gradients in one element-per-workgroup pass (cofactors from the
// coordinates, no stored metric terms) and reduces |grad u|^2 B and
// chi (u+v+w) B with the deterministic reduction tree.
#include "internal.h"
#include "nekoamd.h"
#include "reduce.h"

namespace nekoamd {

constexpr int kTB = 256;

// Scratch of the allocation-free objective (nekoamd_dissipation_objective_
// noalloc): two partials per element and the two sums, for at most
// kObjMaxElements elements (64^3): 4 MiB + 16 B in every device's global
// memory, held until the process ends.  Host code never passes `scratch`
// itself (in host code it names a host-side stand-in); it passes the device
// address that objective_scratch_setup obtained for the current device.
// Every call on a device overwrites the same array; calls that overlap on
// one device (from several host threads) race.  Nothing in it has to start
// at zero: the objective kernels write every partial and both sums.
constexpr int kObjMaxElements = 1 << 18;
constexpr int kObjScratchLen = 2 * kObjMaxElements + 2;
static __device__ __align__(256) double scratch[kObjScratchLen];
constexpr int kScratchDevices = 64;
static double* g_scratch_dev[kScratchDevices] = {};

int objective_scratch_setup(int device) {
  if (device < 0 || device >= kScratchDevices)
    return fail_arg("objective_scratch_setup", "device index out of range");
  if (g_scratch_dev[device]) return kOk;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return fail_hip("objective_scratch_setup", e);
  if (p == nullptr)
    return fail_arg("objective_scratch_setup", "no device address for the scratch");
  g_scratch_dev[device] = (double*)p;
  return kOk;
}

template <typename F>
__global__ void __launch_bounds__(kTB) tmap_k(int64_t n, F f) {
  for (int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kTB)
    f(i);
}

template <typename F>
static int tmap(int64_t n, F f) {
  if (n <= 0) return kOk;
  const int g = nk_grid(n, kTB, kMapGridCap);
  hipLaunchKernelGGL((tmap_k<F>), dim3(g), dim3(kTB), 0,
                     stream(), n, f);
  NK_LAUNCHED();
  return kOk;
}

// Cofactor rows of the coordinate Jacobian and J, from the r/s/t
// derivatives of x, y, z: c_r = x_s x x_t, c_s = x_t x x_r, c_t = x_r x x_s.
struct Cof {
  double r[3], s[3], t[3], J;
};

__device__ __forceinline__ Cof cofactors(const double xr[3], const double xs[3],
                                         const double xt[3]) {
  Cof c;
  c.r[0] = xs[1] * xt[2] - xs[2] * xt[1];
  c.r[1] = xs[2] * xt[0] - xs[0] * xt[2];
  c.r[2] = xs[0] * xt[1] - xs[1] * xt[0];
  c.s[0] = xt[1] * xr[2] - xt[2] * xr[1];
  c.s[1] = xt[2] * xr[0] - xt[0] * xr[2];
  c.s[2] = xt[0] * xr[1] - xt[1] * xr[0];
  c.t[0] = xr[1] * xs[2] - xr[2] * xs[1];
  c.t[1] = xr[2] * xs[0] - xr[0] * xs[2];
  c.t[2] = xr[0] * xs[1] - xr[1] * xs[0];
  c.J = xr[0] * c.r[0] + xr[1] * c.r[1] + xr[2] * c.r[2];
  return c;
}

// Element-local gradient of NF fields (and, when OBJ, the objective
// partials).  One workgroup per element, thread per (i,j) column.
template <int LX, int NF, bool OBJ>
__global__ void __launch_bounds__(LX * LX)
    grad_k(const double* __restrict__ xc, const double* __restrict__ yc,
           const double* __restrict__ zc, const double* __restrict__ dx,
           const double* __restrict__ f0, const double* __restrict__ f1,
           const double* __restrict__ f2, double* __restrict__ gx,
           double* __restrict__ gy, double* __restrict__ gz,
           const double* __restrict__ chi, const double* __restrict__ bm,
           double* __restrict__ part) {
  constexpr int NS = LX * LX, NP = NS * LX;
  __shared__ double s[3 + NF][NP];
  const int e = blockIdx.x, t = threadIdx.x, i = t % LX, j = t / LX;
  const int64_t eb = (int64_t)e * NP + t;
  const double* src[6] = {xc, yc, zc, f0, f1, f2};
#pragma unroll
  for (int a = 0; a < 3 + NF; ++a)
#pragma unroll
    for (int k = 0; k < LX; ++k) s[a][k * NS + t] = src[a][eb + k * NS];
  __syncthreads();
  double diss = 0.0, lube = 0.0;
#pragma unroll
  for (int k = 0; k < LX; ++k) {
    double dr[3 + NF], ds[3 + NF], dt[3 + NF];
#pragma unroll
    for (int a = 0; a < 3 + NF; ++a) {
      double r = 0.0, q = 0.0, w = 0.0;
#pragma unroll
      for (int m = 0; m < LX; ++m) {
        r += dx[i * LX + m] * s[a][k * NS + j * LX + m];
        q += dx[j * LX + m] * s[a][k * NS + m * LX + i];
        w += dx[k * LX + m] * s[a][m * NS + t];
      }
      dr[a] = r;
      ds[a] = q;
      dt[a] = w;
    }
    const double xr[3] = {dr[0], dr[1], dr[2]};
    const double xs[3] = {ds[0], ds[1], ds[2]};
    const double xt[3] = {dt[0], dt[1], dt[2]};
    const Cof c = cofactors(xr, xs, xt);
    const int64_t g = eb + k * NS;
    double sq = 0.0;
#pragma unroll
    for (int a = 0; a < NF; ++a) {
      const double u0 = (c.r[0] * dr[3 + a] + c.s[0] * ds[3 + a] + c.t[0] * dt[3 + a]) / c.J;
      const double u1 = (c.r[1] * dr[3 + a] + c.s[1] * ds[3 + a] + c.t[1] * dt[3 + a]) / c.J;
      const double u2 = (c.r[2] * dr[3 + a] + c.s[2] * ds[3 + a] + c.t[2] * dt[3 + a]) / c.J;
      if constexpr (!OBJ) {
        gx[g] = u0;
        gy[g] = u1;
        gz[g] = u2;
      } else {
        sq += u0 * u0 + u1 * u1 + u2 * u2;
      }
    }
    if constexpr (OBJ) {
      const double b = bm[g];
      diss += sq * b;
      double su = s[3][k * NS + t];
      if constexpr (NF > 1) su += s[4][k * NS + t];
      if constexpr (NF > 2) su += s[5][k * NS + t];
      lube += chi[g] * su * b;
    }
  }
  if constexpr (OBJ) {
    __shared__ double sd[2][NS];
    sd[0][t] = diss;
    sd[1][t] = lube;
    __syncthreads();
    if (t == 0) {
      double a = 0.0, b = 0.0;
      for (int q = 0; q < NS; ++q) {
        a += sd[0][q];
        b += sd[1][q];
      }
      part[e] = a;
      part[(int64_t)gridDim.x + e] = b;
    }
  }
}

__global__ void __launch_bounds__(kFinBlock)
    fin2_k(const double* __restrict__ part, int64_t np, double* out) {
  __shared__ double sh[kFinBlock / 64];
  const double a = sum_partials<kFinBlock>(part, np, sh);
  const double b = sum_partials<kFinBlock>(part + np, np, sh);
  if (threadIdx.x == 0) {
    out[0] = a;
    out[1] = b;
  }
}

template <int NF, bool OBJ>
static int grad_dispatch(int lx, int nelv, const double* x, const double* y,
                         const double* z, const double* dx, const double* f0,
                         const double* f1, const double* f2, double* gx,
                         double* gy, double* gz, const double* chi,
                         const double* B, double* part) {
#define NK_GRAD_CASE(L)                                                      \
  case L:                                                                    \
    hipLaunchKernelGGL((grad_k<L, NF, OBJ>), dim3(nelv), dim3(L * L), 0,     \
                       stream(), x, y, z, dx, f0, f1, f2, gx, gy, gz, chi, B, \
                       part);                                                \
    break;
  switch (lx) {
    NK_GRAD_CASE(2)
    NK_GRAD_CASE(3)
    NK_GRAD_CASE(4)
    NK_GRAD_CASE(5)
    NK_GRAD_CASE(6)
    NK_GRAD_CASE(7)
    NK_GRAD_CASE(8)
    default:
      return fail_arg("grad", "lx must be in [2, 8]");
  }
#undef NK_GRAD_CASE
  NK_LAUNCHED();
  return kOk;
}

// The objective's two sums (dissipation, lube) into host memory hres: the
// gradient pass writes 2 nelv partials into part, fin2_k sums them into
// part[2 nelv], part[2 nelv + 1], which are copied back on the library
// stream.  part holds 2 nelv + 2 doubles.
static int objective_sums(double* hres, double* part, const void* u,
                          const void* v, const void* w, const void* chi,
                          const void* B, const void* x, const void* y,
                          const void* z, const void* dx, int lx, int ne) {
  int rc = grad_dispatch<3, true>(lx, ne, (const double*)x, (const double*)y,
                                  (const double*)z, (const double*)dx,
                                  (const double*)u, (const double*)v,
                                  (const double*)w, nullptr, nullptr, nullptr,
                                  (const double*)chi, (const double*)B, part);
  if (rc) return rc;
  hipLaunchKernelGGL(fin2_k, dim3(1), dim3(kFinBlock), 0, stream(), part,
                     (int64_t)ne, part + 2 * (size_t)ne);
  NK_LAUNCHED();
  NK_HIP(hipMemcpyAsync(hres, part + 2 * (size_t)ne, 2 * sizeof(double),
                        hipMemcpyDeviceToHost, stream()));
  NK_HIP(hipStreamSynchronize(stream()));
  return kOk;
}

// out[0] = (diss + 0.5 K lube) obj_scale (lube only if if_lube),
// out[1] = diss, out[2] = lube.
static void objective_finish(double* out, const double* hres, double K,
                             double obj_scale, int if_lube) {
  const double diss = hres[0], lube = hres[1];
  double F = diss;
  if (if_lube) F = diss + 0.5 * K * lube;
  out[0] = F * obj_scale;
  out[1] = diss;
  out[2] = if_lube ? lube : 0.0;
}

}  // namespace nekoamd

using namespace nekoamd;

#define NK_TN(n) \
  NK_INIT();     \
  NK_REQUIRE(n && *n >= 0, "bad length")

extern "C" {

int nekoamd_brinkman(void* fx, void* fy, void* fz, const void* u,
                     const void* v, const void* w, const void* chi,
                     const int64_t* n) {
  NK_TN(n);
  double *Fx = (double*)fx, *Fy = (double*)fy, *Fz = (double*)fz;
  const double *U = (const double*)u, *V = (const double*)v,
               *W = (const double*)w, *X = (const double*)chi;
  return tmap(*n, [=] __device__(int64_t i) {
    const double c = X[i];
    Fx[i] = Fx[i] - U[i] * c;
    Fy[i] = Fy[i] - V[i] * c;
    Fz[i] = Fz[i] - W[i] * c;
  });
}

int nekoamd_brinkman_mass(void* fx, void* fy, void* fz, const void* u,
                          const void* v, const void* w, const void* chi,
                          const void* B, const int64_t* n) {
  NK_TN(n);
  double *Fx = (double*)fx, *Fy = (double*)fy, *Fz = (double*)fz;
  const double *U = (const double*)u, *V = (const double*)v,
               *W = (const double*)w, *X = (const double*)chi,
               *M = (const double*)B;
  return tmap(*n, [=] __device__(int64_t i) {
    const double c = X[i] * M[i];
    Fx[i] = Fx[i] - U[i] * c;
    Fy[i] = Fy[i] - V[i] * c;
    Fz[i] = Fz[i] - W[i] * c;
  });
}

int nekoamd_update_permeability(void* res, const void* rho,
                                const double* perm_0, const double* perm_1,
                                const double* penalty, const int64_t* n) {
  NK_TN(n);
  NK_REQUIRE(perm_0 && perm_1 && penalty, "null scalar");
  double* R = (double*)res;
  const double* P = (const double*)rho;
  const double q = *penalty, k1 = *perm_1;
  const double constant = (*perm_0 - *perm_1) * (q + 1.0);
  return tmap(*n, [=] __device__(int64_t i) {
    double r = q + P[i];
    r = 1.0 / r;
    r = r * P[i];
    r = constant * r;
    r = r + k1;
    R[i] = -1.0 * r;
  });
}

int nekoamd_ramp_fwd(void* out, const void* in, const double* f_min,
                     const double* f_max, const double* q,
                     const int* convex_up, const int64_t* n) {
  NK_TN(n);
  NK_REQUIRE(f_min && f_max && q && convex_up, "null scalar");
  double* O = (double*)out;
  const double* X = (const double*)in;
  const double lo = *f_min, hi = *f_max, Q = *q;
  if (*convex_up)
    return tmap(*n, [=] __device__(int64_t i) {
      const double x = X[i];
      O[i] = lo + (hi - lo) * x * (1.0 + Q) / (x + Q);
    });
  return tmap(*n, [=] __device__(int64_t i) {
    const double x = X[i];
    O[i] = lo + (hi - lo) * x / (1.0 + Q * (1.0 - x));
  });
}

int nekoamd_ramp_bwd(void* df_din, const void* df_dout, const void* in,
                     const double* f_min, const double* f_max, const double* q,
                     const int* convex_up, const int64_t* n) {
  NK_TN(n);
  NK_REQUIRE(f_min && f_max && q && convex_up, "null scalar");
  double* O = (double*)df_din;
  const double *G = (const double*)df_dout, *X = (const double*)in;
  const double lo = *f_min, hi = *f_max, Q = *q;
  if (*convex_up)
    return tmap(*n, [=] __device__(int64_t i) {
      const double a = X[i] + Q;
      O[i] = (hi - lo) * (Q + 1.0) / (a * a) * G[i];
    });
  return tmap(*n, [=] __device__(int64_t i) {
    const double a = 1.0 - Q * (X[i] - 1.0);
    O[i] = (hi - lo) * (Q + 1.0) / (a * a) * G[i];
  });
}

int nekoamd_linear_fwd(void* out, const void* in, const double* f_min,
                       const double* f_max, const int64_t* n) {
  NK_TN(n);
  NK_REQUIRE(f_min && f_max, "null scalar");
  double* O = (double*)out;
  const double* X = (const double*)in;
  const double lo = *f_min, s = *f_max - *f_min;
  return tmap(*n, [=] __device__(int64_t i) { O[i] = s * X[i] + lo; });
}

int nekoamd_linear_bwd(void* df_din, const void* df_dout, const double* f_min,
                       const double* f_max, const int64_t* n) {
  NK_TN(n);
  NK_REQUIRE(f_min && f_max, "null scalar");
  double* O = (double*)df_din;
  const double* G = (const double*)df_dout;
  const double s = *f_max - *f_min;
  return tmap(*n, [=] __device__(int64_t i) { O[i] = s * G[i]; });
}

int nekoamd_dissipation_sensitivity(void* s, const void* u, const void* v,
                                    const void* w, const void* ua,
                                    const void* va, const void* wa,
                                    const double* K, const double* obj_scale,
                                    const int* if_lube, const int64_t* n) {
  NK_TN(n);
  NK_REQUIRE(K && obj_scale && if_lube, "null scalar");
  double* S = (double*)s;
  const double *U = (const double*)u, *V = (const double*)v,
               *W = (const double*)w, *Ua = (const double*)ua,
               *Va = (const double*)va, *Wa = (const double*)wa;
  const double kk = *K * *obj_scale;
  const bool lube = *if_lube != 0;
  return tmap(*n, [=] __device__(int64_t i) {
    const double u0 = U[i], v0 = V[i], w0 = W[i];
    double a = u0 * Ua[i];
    a = a + v0 * Va[i];
    a = a + w0 * Wa[i];
    a = -1.0 * a;
    if (lube) {
      double l = u0 * u0;
      l = l + v0 * v0;
      l = l + w0 * w0;
      a = a + kk * l;
    }
    S[i] = a;
  });
}

int nekoamd_adjoint_lube(void* fx, void* fy, void* fz, const void* u,
                         const void* v, const void* w, const void* chi,
                         const double* K, const int64_t* n) {
  NK_TN(n);
  NK_REQUIRE(K, "null scalar");
  double *Fx = (double*)fx, *Fy = (double*)fy, *Fz = (double*)fz;
  const double *U = (const double*)u, *V = (const double*)v,
               *W = (const double*)w, *X = (const double*)chi;
  const double k = *K;
  return tmap(*n, [=] __device__(int64_t i) {
    const double c = X[i] * k;
    Fx[i] = Fx[i] + U[i] * c;
    Fy[i] = Fy[i] + V[i] * c;
    Fz[i] = Fz[i] + W[i] * c;
  });
}

int nekoamd_grad(void* ux, void* uy, void* uz, const void* u, const void* x,
                 const void* y, const void* z, const void* dx, const int* lx,
                 const int* nelv) {
  NK_INIT();
  NK_REQUIRE(ux && uy && uz && u && x && y && z && dx && lx && nelv,
             "null argument");
  if (*nelv <= 0) return kOk;
  return grad_dispatch<1, false>(*lx, *nelv, (const double*)x, (const double*)y,
                                 (const double*)z, (const double*)dx,
                                 (const double*)u, nullptr, nullptr,
                                 (double*)ux, (double*)uy, (double*)uz,
                                 nullptr, nullptr, nullptr);
}

// out[0] = (diss + 0.5 K lube) obj_scale (lube only if *if_lube),
// out[1] = diss, out[2] = lube.
int nekoamd_dissipation_objective(double* out, const void* u, const void* v,
                                  const void* w, const void* chi,
                                  const void* B, const void* x, const void* y,
                                  const void* z, const void* dx, const int* lx,
                                  const int* nelv, const double* K,
                                  const double* obj_scale, const int* if_lube) {
  NK_INIT();
  NK_REQUIRE(out && u && v && w && chi && B && x && y && z && dx && lx &&
                 nelv && K && obj_scale && if_lube,
             "null argument");
  const int ne = *nelv;
  if (ne <= 0) {
    out[0] = out[1] = out[2] = 0.0;
    return kOk;
  }
  double* part = nullptr;
  double* hres = nullptr;
  NK_HIP(hipMalloc((void**)&part, sizeof(double) * (2 * (size_t)ne + 2)));
  hipError_t e = hipHostMalloc((void**)&hres, 2 * sizeof(double),
                               hipHostMallocDefault);
  int rc = e == hipSuccess ? kOk : fail_hip(__func__, e);
  if (rc == kOk)
    rc = objective_sums(hres, part, u, v, w, chi, B, x, y, z, dx, *lx, ne);
  if (rc == kOk) objective_finish(out, hres, *K, *obj_scale, *if_lube);
  if (hres) (void)hipHostFree(hres);
  (void)hipFree(part);
  return rc;
}

// The same objective with its partials in the library's static scratch
// (at most kObjMaxElements elements): no device or host allocation.
int nekoamd_dissipation_objective_noalloc(
    double* out, const void* u, const void* v, const void* w, const void* chi,
    const void* B, const void* x, const void* y, const void* z,
    const void* dx, const int* lx, const int* nelv, const double* K,
    const double* obj_scale, const int* if_lube) {
  NK_INIT();
  NK_REQUIRE(out && u && v && w && chi && B && x && y && z && dx && lx &&
                 nelv && K && obj_scale && if_lube,
             "null argument");
  const int ne = *nelv;
  NK_REQUIRE(ne <= kObjMaxElements,
             "nelv exceeds the allocation-free maximum of 262144 elements");
  if (ne <= 0) {
    out[0] = out[1] = out[2] = 0.0;
    return kOk;
  }
  const int dev = current_device();
  double* part =
      dev >= 0 && dev < kScratchDevices ? g_scratch_dev[dev] : nullptr;
  if (!part) {
    set_error(__func__, "no scratch address for this device");
    return kErrState;
  }
  double hres[2] = {0.0, 0.0};
  const int rc =
      objective_sums(hres, part, u, v, w, chi, B, x, y, z, dx, *lx, ne);
  if (rc == kOk) objective_finish(out, hres, *K, *obj_scale, *if_lube);
  return rc;
}


int nekoamd_volume_constraint(double* out, const void* rho, const void* B,
                              const double* volume_domain, const double* v_max,
                              const int* is_max, const int64_t* n) {
  NK_TN(n);
  NK_REQUIRE(out && volume_domain && v_max && is_max, "null argument");
  double s = 0.0;
  int rc = nekoamd_glsc2(&s, rho, B, n);
  if (rc) return rc;
  const double V = s / *volume_domain;
  out[1] = V;
  out[0] = *is_max ? V - *v_max : -V + *v_max;
  return kOk;
}

int nekoamd_volume_sensitivity(void* s, const double* volume_domain,
                               const int* is_max, const int64_t* n) {
  NK_TN(n);
  NK_REQUIRE(volume_domain && is_max, "null argument");
  const double c = *is_max ? 1.0 / *volume_domain : -1.0 / *volume_domain;
  return nekoamd_cfill(s, &c, n);
}

}  // extern "C"
