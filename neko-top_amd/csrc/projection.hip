// projection.hip -- solution projection for successive right-hand sides
// (Fischer, CMAME 163, 1998), the method of Neko's projection_t.
//
// Replaces proj_prs / proj_u / proj_v / proj_w of the reference
// (sources/adjoint/adjoint_pnpn.f90:102-105; init :365-373; pre_solving
// :729-730, :777-779; post_solving :739-740, :788-793).
//
// The basis lives in HBM as two L x n FP64 blocks, xx (A-orthonormal
// previous solutions) and bb = A xx.  Per solve:
//   pre  : alpha_i = <b, xx_i>_mult (one pass per basis vector, fixed-order
//          partials), then one fused pass xbar = sum alpha_i xx_i,
//          b -= sum alpha_i bb_i
//   post : x += xbar; xx_k = x; bb_k = mask(QQ^T A x) (ax_helm + gs + bc);
//          two passes of classical Gram-Schmidt in the A-inner product;
//          keep the vector scaled to A-norm 1 unless it is dependent.
// Coefficients stay on the device; the host reads two norms per post.
// Algorithmic bytes: pre (2m + 3) x 8 B/point, post ~ (4m + 12) x 8 B/point
// plus one operator application: a few CG iterations' worth per solve.
#include "internal.h"
#include "nekoamd.h"
#include "reduce.h"

namespace nekoamd {

constexpr double kProjTol = 1e-7;
constexpr int kDotGrid = 512;  // partials per inner product

struct ProjState {
  int64_t n = 0;
  int L = 0;
  int m = 0;
  bool xbar_valid = false;
  double* xx = nullptr;    // L * n
  double* bb = nullptr;    // L * n
  double* xbar = nullptr;  // n
  double* coef = nullptr;  // NEKOAMD_PROJ_MAX_L + 2 (alpha..., then scratch)
  double* part = nullptr;  // (NEKOAMD_PROJ_MAX_L + 2) * kDotGrid
  double* host = nullptr;  // pinned, 4
};

// part[i * G + blk] = partial of sum_p mult a V_i over block blk's points.
__global__ void __launch_bounds__(256)
    proj_mdot_k(int64_t n, const double* __restrict__ a,
                const double* __restrict__ V, int64_t ldv,
                const double* __restrict__ mult, double* __restrict__ part) {
  __shared__ double sh[256 / 64];
  const int i = blockIdx.y;
  const double* v = V + (int64_t)i * ldv;
  double s = 0.0;
  for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * 256)
    s += mult[p] * a[p] * v[p];
  s = block_sum<256>(s, sh);
  if (threadIdx.x == 0) part[(int64_t)i * gridDim.x + blockIdx.x] = s;
}

// out[i] = sum of the G partials of vector i (one block per i).
__global__ void __launch_bounds__(kFinBlock)
    proj_mdot_fin_k(const double* __restrict__ part, int G,
                    double* __restrict__ out) {
  __shared__ double sh[kFinBlock / 64];
  const double s =
      sum_partials<kFinBlock>(part + (int64_t)blockIdx.x * G, G, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// x_out = sum alpha_i xx_i (assign) or x_out -= sum alpha_i xx_i;
// b -= sum alpha_i bb_i.
__global__ void __launch_bounds__(256)
    proj_lincomb_k(int64_t n, int m, const double* __restrict__ alpha,
                   const double* __restrict__ xx,
                   const double* __restrict__ bb, double* __restrict__ xout,
                   double* __restrict__ b, int assign) {
  __shared__ double sa[NEKOAMD_PROJ_MAX_L];
  if (threadIdx.x < m) sa[threadIdx.x] = alpha[threadIdx.x];
  __syncthreads();
  for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * 256) {
    double sx = 0.0, sb = 0.0;
    for (int i = 0; i < m; ++i) {
      sx += sa[i] * xx[(int64_t)i * n + p];
      sb += sa[i] * bb[(int64_t)i * n + p];
    }
    xout[p] = assign ? sx : xout[p] - sx;
    b[p] -= sb;
  }
}

__global__ void __launch_bounds__(256)
    proj_axpy_k(int64_t n, double* __restrict__ y, const double* __restrict__ x,
                double a) {
  for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * 256)
    y[p] += a * x[p];
}

__global__ void __launch_bounds__(256)
    proj_scale2_k(int64_t n, double* __restrict__ u, double* __restrict__ v,
                  double s) {
  for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * 256) {
    u[p] *= s;
    v[p] *= s;
  }
}

// gamma = <q,b>/<q,Aq> (0 if <q,Aq> <= 0); xbar = [xbar +] gamma q;
// b -= gamma Aq.  dots = {<q,b>, <q,Aq>}; block 0 stores gamma in dots[2].
__global__ void __launch_bounds__(256)
    proj_augment_k(int64_t n, const double* __restrict__ q,
                   const double* __restrict__ aq, double* __restrict__ xbar,
                   double* __restrict__ b, double* __restrict__ dots,
                   int assign) {
  const double qaq = dots[1];
  const double g = qaq > 0.0 ? dots[0] / qaq : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) dots[2] = g;
  for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * 256) {
    xbar[p] = (assign ? 0.0 : xbar[p]) + g * q[p];
    b[p] -= g * aq[p];
  }
}

static ProjState* as_proj(void* p) { return (ProjState*)p; }

// out[0..m) = <a, V_i>_mult for the m vectors V_i = V + i * ldv.
static int mdot(ProjState* s, const double* a, const double* V, int64_t ldv,
                int m, const double* mult, double* out) {
  if (m <= 0) return kOk;
  const int G = nk_grid(s->n, 256, kDotGrid);
  hipLaunchKernelGGL(proj_mdot_k, dim3(G, m), dim3(256), 0, stream(), s->n, a,
                     V, ldv, mult, s->part);
  NK_LAUNCHED();
  hipLaunchKernelGGL(proj_mdot_fin_k, dim3(m), dim3(kFinBlock), 0, stream(),
                     s->part, G, out);
  NK_LAUNCHED();
  return kOk;
}

}  // namespace nekoamd

using namespace nekoamd;

extern "C" {

int nekoamd_proj_create(void** proj, const int64_t* n, const int* L) {
  NK_INIT();
  NK_REQUIRE(proj && n && L && *n >= 0, "bad arguments");
  NK_REQUIRE(*L >= 1 && *L <= NEKOAMD_PROJ_MAX_L,
             "L must be in [1, NEKOAMD_PROJ_MAX_L]");
  ProjState* s = new ProjState();
  s->n = *n;
  s->L = *L;
  const int64_t N = s->n > 0 ? s->n : 1;
  hipError_t e = hipMalloc((void**)&s->xx, sizeof(double) * N * s->L);
  if (e == hipSuccess)
    e = hipMalloc((void**)&s->bb, sizeof(double) * N * s->L);
  if (e == hipSuccess) e = hipMalloc((void**)&s->xbar, sizeof(double) * N);
  if (e == hipSuccess)
    e = hipMalloc((void**)&s->coef,
                  sizeof(double) * (NEKOAMD_PROJ_MAX_L + 4));
  if (e == hipSuccess)
    e = hipMalloc((void**)&s->part,
                  sizeof(double) * (NEKOAMD_PROJ_MAX_L + 2) * kDotGrid);
  if (e == hipSuccess)
    e = hipHostMalloc((void**)&s->host, 4 * sizeof(double),
                      hipHostMallocDefault);
  if (e != hipSuccess) {
    nekoamd_proj_free(s);
    return fail_hip(__func__, e);
  }
  *proj = s;
  return kOk;
}

int nekoamd_proj_free(void* proj) {
  ProjState* s = as_proj(proj);
  if (!s) return kOk;
  if (initialised()) NK_HIP(hipStreamSynchronize(stream()));
  void* dev[] = {s->xx, s->bb, s->xbar, s->coef, s->part};
  for (void* q : dev)
    if (q) NK_HIP(hipFree(q));
  if (s->host) NK_HIP(hipHostFree(s->host));
  delete s;
  return kOk;
}

int nekoamd_proj_clear(void* proj) {
  ProjState* s = as_proj(proj);
  if (!s) return fail_arg(__func__, "null projection");
  s->m = 0;
  s->xbar_valid = false;
  return kOk;
}

int nekoamd_proj_size(void* proj, int* m) {
  ProjState* s = as_proj(proj);
  if (!s || !m) return fail_arg(__func__, "null argument");
  *m = s->m;
  return kOk;
}

int nekoamd_proj_pre(void* proj, void* b, const void* mult) {
  NK_INIT();
  ProjState* s = as_proj(proj);
  NK_REQUIRE(s && (s->n == 0 || (b && mult)), "null argument");
  s->xbar_valid = false;
  if (s->m == 0 || s->n == 0) return kOk;
  int rc = mdot(s, (const double*)b, s->xx, s->n, s->m, (const double*)mult,
                s->coef);
  if (rc) return rc;
  hipLaunchKernelGGL(proj_lincomb_k, dim3(nk_grid(s->n, 256)), dim3(256), 0,
                     stream(), s->n, s->m, s->coef, s->xx, s->bb, s->xbar,
                     (double*)b, 1);
  NK_LAUNCHED();
  s->xbar_valid = true;
  return kOk;
}

int nekoamd_proj_augment(void* proj, void* b, const void* q, const void* aq,
                         const void* mult, double* gamma) {
  NK_INIT();
  ProjState* s = as_proj(proj);
  NK_REQUIRE(s && gamma && (s->n == 0 || (b && q && aq && mult)),
             "null argument");
  if (s->n == 0) {
    *gamma = 0.0;
    return kOk;
  }
  double* dots = s->coef + NEKOAMD_PROJ_MAX_L;
  int rc = mdot(s, (const double*)q, (const double*)b, 0, 1,
                (const double*)mult, dots);
  if (!rc)
    rc = mdot(s, (const double*)q, (const double*)aq, 0, 1,
              (const double*)mult, dots + 1);
  if (rc) return rc;
  hipLaunchKernelGGL(proj_augment_k, dim3(nk_grid(s->n, 256)), dim3(256), 0,
                     stream(), s->n, (const double*)q, (const double*)aq,
                     s->xbar, (double*)b, dots, s->xbar_valid ? 0 : 1);
  NK_LAUNCHED();
  s->xbar_valid = true;
  NK_HIP(hipMemcpyAsync(s->host, dots + 2, sizeof(double),
                        hipMemcpyDeviceToHost, stream()));
  NK_HIP(hipStreamSynchronize(stream()));
  *gamma = s->host[0];
  return kOk;
}

int nekoamd_proj_post(void* proj, void* x, const void* mult, void* gs,
                      const void* mask, const int64_t* nmask, const void* dx,
                      const void* h1, const void* h2, const void* g11,
                      const void* g12, const void* g13, const void* g22,
                      const void* g23, const void* g33, const void* B,
                      const int* lx, const int* nelv, const int* ifh2) {
  NK_INIT();
  ProjState* s = as_proj(proj);
  NK_REQUIRE(s && nmask && lx && nelv && ifh2, "null argument");
  NK_REQUIRE((int64_t)(*nelv) * (*lx) * (*lx) * (*lx) == s->n,
             "nelv * lx^3 does not match the projection size");
  NK_REQUIRE(s->n == 0 || (x && mult && gs && dx && h1 && g11 && g12 && g13 &&
                           g22 && g23 && g33),
             "null field");
  NK_REQUIRE(*nmask >= 0 && (*nmask == 0 || mask), "bad mask");
  NK_REQUIRE(!*ifh2 || (h2 && B), "ifh2 set but h2 or B is null");
  if (s->n == 0) return kOk;
  const int64_t n = s->n;
  const int ge = nk_grid(n, 256);
  double* X = (double*)x;
  const double* M = (const double*)mult;
  if (s->xbar_valid) {
    hipLaunchKernelGGL(proj_axpy_k, dim3(ge), dim3(256), 0, stream(), n, X,
                       s->xbar, 1.0);
    NK_LAUNCHED();
  }
  s->xbar_valid = false;
  if (s->m == s->L) s->m = 0;
  const int k = s->m;
  double* xk = s->xx + (int64_t)k * n;
  double* bk = s->bb + (int64_t)k * n;
  NK_HIP(hipMemcpyAsync(xk, X, sizeof(double) * n, hipMemcpyDeviceToDevice,
                        stream()));
  int rc = ax_helm_launch<double, double>(
      bk, xk, (const double*)dx, (const double*)h1, (const double*)h2,
      (const double*)g11, (const double*)g12, (const double*)g13,
      (const double*)g22, (const double*)g23, (const double*)g33,
      (const double*)B, *lx, *nelv, *ifh2 != 0, nullptr);
  if (!rc) rc = gs_launch<double>(gs, bk, NEKOAMD_GS_ADD, nullptr);
  if (!rc)
    rc = bc_zero_launch<double>(bk, (const int32_t*)mask, *nmask, nullptr);
  if (rc) return rc;
  double* nrm = s->coef + NEKOAMD_PROJ_MAX_L;  // [0]: before, [1]: after
  rc = mdot(s, xk, bk, 0, 1, M, nrm);
  if (rc) return rc;
  for (int pass = 0; pass < 2 && k > 0; ++pass) {
    rc = mdot(s, xk, s->bb, n, k, M, s->coef);
    if (rc) return rc;
    hipLaunchKernelGGL(proj_lincomb_k, dim3(ge), dim3(256), 0, stream(), n, k,
                       s->coef, s->xx, s->bb, xk, bk, 0);
    NK_LAUNCHED();
  }
  rc = mdot(s, xk, bk, 0, 1, M, nrm + 1);
  if (rc) return rc;
  NK_HIP(hipMemcpyAsync(s->host, nrm, 2 * sizeof(double),
                        hipMemcpyDeviceToHost, stream()));
  NK_HIP(hipStreamSynchronize(stream()));
  const double n0 = sqrt(s->host[0] > 0.0 ? s->host[0] : 0.0);
  const double n1 = sqrt(s->host[1] > 0.0 ? s->host[1] : 0.0);
  if (n1 > kProjTol * n0 && n1 > 0.0) {
    hipLaunchKernelGGL(proj_scale2_k, dim3(ge), dim3(256), 0, stream(), n, xk,
                       bk, 1.0 / n1);
    NK_LAUNCHED();
    s->m = k + 1;
  }
  return kOk;
}

}  // extern "C"
