// cg.hip -- Jacobi preconditioner and preconditioned conjugate gradients
// with device-resident scalars.
//
// Replaces Neko's jacobi_t / device_jacobi_t (precon_factory 'jacobi' at
// sources/adjoint/adjoint_scheme.f90:1068-1076; update() at
// sources/adjoint/adjoint_pnpn.f90:732, :781) and ksp_t%solve for 'cg'
// (krylov_solver_factory at adjoint_scheme.f90:1046-1055; solves at
// adjoint_pnpn.f90:734-736 and, per component, :784-786).
//
// One CG iteration is five launches on the library stream:
//   pupdate : p = d r + beta p
//   ax      : w = A_L p, fused per-element partials of w.p          (ax_helm.hip)
//   gs_bc   : w = mask(QQ^T w)                                      (gs.hip)
//   alpha   : pap = sum partials, alpha = rtz / pap       (one block)
//   update  : x += alpha p, r -= alpha w, partials of r.r.mult and
//             r.(d r).mult, then (one block) beta, ||r||, stop flag
// alpha, beta, the norm and the stop flag live in device memory, so the host
// never waits inside an iteration; once the stop flag is up every later
// kernel returns immediately, and the host polls the flag every few
// iterations.  Iteration counts and stopping follow the reference's CG:
// norm sqrt(<r,r>_mult)/sqrt(volume), checked after each iteration.
//
// The fused pap = sum_local w.p equals the reference's <QQ^T w, p>_mult
// because p is continuous and zero on the Dirichlet mask (r starts as the
// gs-assembled, masked right-hand side and stays so).
#include "internal.h"
#include "nekoamd.h"
#include "reduce.h"

namespace nekoamd {

struct CgScal {
  double rtz1, rtz2, pap, alpha, beta, rtr, rnorm, res_start, norm_fac, abstol;
  int iter, max_iter, done, pad;
};

struct CgState {
  int64_t n = 0;
  int nelv = 0;
  int grid = 1;
  CgScal* sc = nullptr;      // device
  double* part_ax = nullptr; // nelv
  double* part = nullptr;    // 2 * kRedGrid
  CgScal* host = nullptr;    // pinned, 2 slots
};

// ---------------------------------------------------------------------------
template <int LX>
__global__ void __launch_bounds__(LX * LX)
    jacobi_diag_k(double* __restrict__ d, const double* __restrict__ dx,
                  const double* __restrict__ g11,
                  const double* __restrict__ g22,
                  const double* __restrict__ g33,
                  const double* __restrict__ h1, const double* __restrict__ h2,
                  const double* __restrict__ bm, int ifh2) {
  constexpr int NS = LX * LX, NP = NS * LX;
  __shared__ double s11[NP];
  __shared__ double s22[NP];
  const int e = blockIdx.x, t = threadIdx.x, i = t % LX, j = t / LX;
  const int64_t eb = (int64_t)e * NP + t;
  double c33[LX];
#pragma unroll
  for (int k = 0; k < LX; ++k) {
    s11[k * NS + t] = g11[eb + k * NS];
    s22[k * NS + t] = g22[eb + k * NS];
    c33[k] = g33[eb + k * NS];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < LX; ++k) {
    double a = 0.0;
#pragma unroll
    for (int l = 0; l < LX; ++l) {
      const double D = dx[l * LX + i];
      a += s11[k * NS + j * LX + l] * D * D;
    }
#pragma unroll
    for (int l = 0; l < LX; ++l) {
      const double D = dx[l * LX + j];
      a += s22[k * NS + l * LX + i] * D * D;
    }
#pragma unroll
    for (int l = 0; l < LX; ++l) {
      const double D = dx[l * LX + k];
      a += c33[l] * D * D;
    }
    const int64_t g = eb + k * NS;
    a = a * h1[g];
    if (ifh2) a = a + h2[g] * bm[g];
    d[g] = a;
  }
}

// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kRedBlock)
    cg_start_k(int64_t n, double* __restrict__ x, double* __restrict__ r,
               const double* __restrict__ f, const double* __restrict__ mult,
               const double* __restrict__ d, double* __restrict__ part) {
  __shared__ double sh[kRedBlock / 64];
  double a = 0.0, b = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRedBlock) {
    const double ri = f[i];
    x[i] = 0.0;
    r[i] = ri;
    const double rm = ri * mult[i];
    a += rm * ri;
    b += rm * (ri * d[i]);
  }
  a = block_sum<kRedBlock>(a, sh);
  b = block_sum<kRedBlock>(b, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[kRedGrid + blockIdx.x] = b;
  }
}

__global__ void __launch_bounds__(kFinBlock)
    cg_start_fin_k(CgScal* __restrict__ sc, const double* __restrict__ part,
                   int np, double volume, double abstol, int max_iter) {
  __shared__ double sh[kFinBlock / 64];
  const double rtr = sum_partials<kFinBlock>(part, np, sh);
  const double rtz = sum_partials<kFinBlock>(part + kRedGrid, np, sh);
  if (threadIdx.x == 0) {
    const double nf = 1.0 / sqrt(volume);
    const double rn = sqrt(rtr) * nf;
    sc->norm_fac = nf;
    sc->abstol = abstol;
    sc->rtr = rtr;
    sc->rnorm = rn;
    sc->res_start = rn;
    sc->rtz1 = rtz;
    sc->rtz2 = 1.0;
    sc->beta = 0.0;
    sc->alpha = 0.0;
    sc->pap = 0.0;
    sc->iter = 0;
    sc->max_iter = max_iter;
    sc->done = (rn == 0.0 || max_iter <= 0) ? 1 : 0;
  }
}

__global__ void __launch_bounds__(256)
    cg_p_k(int64_t n, double* __restrict__ p, const double* __restrict__ r,
           const double* __restrict__ d, const CgScal* __restrict__ sc) {
  if (sc->done) return;
  const bool first = sc->iter == 0;
  const double beta = sc->beta;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    const double z = r[i] * d[i];
    p[i] = first ? z : z + beta * p[i];
  }
}

__global__ void __launch_bounds__(kFinBlock)
    cg_alpha_k(CgScal* __restrict__ sc, const double* __restrict__ part_ax,
               int nelv) {
  if (sc->done) return;
  __shared__ double sh[kFinBlock / 64];
  const double pap = sum_partials<kFinBlock>(part_ax, nelv, sh);
  if (threadIdx.x == 0) {
    sc->pap = pap;
    sc->alpha = sc->rtz1 / pap;
  }
}

__global__ void __launch_bounds__(kRedBlock)
    cg_update_k(int64_t n, double* __restrict__ x, double* __restrict__ r,
                const double* __restrict__ p, const double* __restrict__ w,
                const double* __restrict__ d, const double* __restrict__ mult,
                const CgScal* __restrict__ sc, double* __restrict__ part) {
  if (sc->done) return;
  __shared__ double sh[kRedBlock / 64];
  const double alpha = sc->alpha;
  double a = 0.0, b = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRedBlock) {
    x[i] = x[i] + alpha * p[i];
    const double ri = r[i] - alpha * w[i];
    r[i] = ri;
    const double rm = ri * mult[i];
    a += rm * ri;
    b += rm * (ri * d[i]);
  }
  a = block_sum<kRedBlock>(a, sh);
  b = block_sum<kRedBlock>(b, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[kRedGrid + blockIdx.x] = b;
  }
}

__global__ void __launch_bounds__(kFinBlock)
    cg_update_fin_k(CgScal* __restrict__ sc, const double* __restrict__ part,
                    int np) {
  if (sc->done) return;
  __shared__ double sh[kFinBlock / 64];
  const double rtr = sum_partials<kFinBlock>(part, np, sh);
  const double rtz = sum_partials<kFinBlock>(part + kRedGrid, np, sh);
  if (threadIdx.x == 0) {
    const int it = sc->iter + 1;
    const double rn = sqrt(rtr) * sc->norm_fac;
    sc->iter = it;
    sc->rtr = rtr;
    sc->rnorm = rn;
    sc->beta = rtz / sc->rtz1;
    sc->rtz2 = sc->rtz1;
    sc->rtz1 = rtz;
    if (rn < sc->abstol || it >= sc->max_iter) sc->done = 1;
  }
}

static CgState* as_cg(void* cg) { return (CgState*)cg; }

}  // namespace nekoamd

using namespace nekoamd;

extern "C" {

int nekoamd_jacobi_update(void* d, void* gs, const void* dx, const void* g11,
                          const void* g22, const void* g33, const void* h1,
                          const void* h2, const void* B, const int* lx,
                          const int* nelv, const int* ifh2) {
  NK_INIT();
  NK_REQUIRE(d && gs && dx && g11 && g22 && g33 && h1 && lx && nelv && ifh2,
             "null argument");
  NK_REQUIRE(!*ifh2 || (h2 && B), "ifh2 set but h2 or B is null");
  if (*nelv <= 0) return kOk;
  const int L = *lx;
  auto* D = (double*)d;
#define NK_JAC_CASE(X)                                                     \
  case X:                                                                  \
    hipLaunchKernelGGL(jacobi_diag_k<X>, dim3(*nelv), dim3(X * X), 0,      \
                       stream(), D, (const double*)dx, (const double*)g11, \
                       (const double*)g22, (const double*)g33,             \
                       (const double*)h1, (const double*)h2,               \
                       (const double*)B, *ifh2);                           \
    break;
  switch (L) {
    NK_JAC_CASE(2)
    NK_JAC_CASE(3)
    NK_JAC_CASE(4)
    NK_JAC_CASE(5)
    NK_JAC_CASE(6)
    NK_JAC_CASE(7)
    NK_JAC_CASE(8)
    NK_JAC_CASE(9)
    NK_JAC_CASE(10)
    default:
      return fail_arg(__func__, "lx must be in [2, 10]");
  }
#undef NK_JAC_CASE
  NK_LAUNCHED();
  int rc = gs_launch<double>(gs, D, NEKOAMD_GS_ADD, nullptr);
  if (rc) return rc;
  const int64_t n = (int64_t)*nelv * L * L * L;
  return nekoamd_invcol1(d, &n);
}

int nekoamd_jacobi_solve(void* z, const void* r, const void* d,
                         const int64_t* n) {
  return nekoamd_col3(z, r, d, n);
}

int nekoamd_cg_create(void** cg, const int64_t* n, const int* nelv) {
  NK_INIT();
  NK_REQUIRE(cg && n && nelv && *n >= 0 && *nelv >= 0, "bad arguments");
  CgState* s = new CgState();
  s->n = *n;
  s->nelv = *nelv;
  s->grid = nk_grid(*n, kRedBlock, kRedGrid);
  hipError_t e = hipMalloc((void**)&s->sc, sizeof(CgScal));
  if (e == hipSuccess)
    e = hipMalloc((void**)&s->part_ax, sizeof(double) * (s->nelv + 1));
  if (e == hipSuccess)
    e = hipMalloc((void**)&s->part, sizeof(double) * 2 * kRedGrid);
  if (e == hipSuccess)
    e = hipHostMalloc((void**)&s->host, 2 * sizeof(CgScal),
                      hipHostMallocDefault);
  if (e != hipSuccess) {
    nekoamd_cg_free(s);
    return fail_hip(__func__, e);
  }
  // a fresh state counts as converged until cg_start
  CgScal init{};
  init.done = 1;
  e = hipMemcpy(s->sc, &init, sizeof(CgScal), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    nekoamd_cg_free(s);
    return fail_hip(__func__, e);
  }
  *cg = s;
  return kOk;
}

int nekoamd_cg_free(void* cg) {
  CgState* s = as_cg(cg);
  if (!s) return kOk;
  if (initialised()) NK_HIP(hipStreamSynchronize(stream()));
  if (s->sc) NK_HIP(hipFree(s->sc));
  if (s->part_ax) NK_HIP(hipFree(s->part_ax));
  if (s->part) NK_HIP(hipFree(s->part));
  if (s->host) NK_HIP(hipHostFree(s->host));
  delete s;
  return kOk;
}

int nekoamd_cg_start(void* cg, void* x, void* r, const void* f,
                     const void* mult, const void* d, const double* volume,
                     const double* abstol, const int* max_iter) {
  NK_INIT();
  CgState* s = as_cg(cg);
  NK_REQUIRE(s && volume && abstol && max_iter, "null argument");
  NK_REQUIRE(s->n == 0 || (x && r && f && mult && d), "null field");
  NK_REQUIRE(*volume > 0.0, "volume must be positive");
  hipLaunchKernelGGL(cg_start_k, dim3(s->grid), dim3(kRedBlock), 0, stream(),
                     s->n, (double*)x, (double*)r, (const double*)f,
                     (const double*)mult, (const double*)d, s->part);
  NK_LAUNCHED();
  hipLaunchKernelGGL(cg_start_fin_k, dim3(1), dim3(kFinBlock), 0, stream(),
                     s->sc, s->part, s->grid, *volume, *abstol, *max_iter);
  NK_LAUNCHED();
  return kOk;
}

int nekoamd_cg_pupdate(void* cg, void* p, const void* r, const void* d) {
  NK_INIT();
  CgState* s = as_cg(cg);
  NK_REQUIRE(s, "null cg");
  hipLaunchKernelGGL(cg_p_k, dim3(nk_grid(s->n, 256)), dim3(256), 0, stream(),
                     s->n, (double*)p, (const double*)r, (const double*)d,
                     s->sc);
  NK_LAUNCHED();
  return kOk;
}

int nekoamd_cg_ax(void* cg, void* w, const void* p, const void* dx,
                  const void* h1, const void* h2, const void* g11,
                  const void* g12, const void* g13, const void* g22,
                  const void* g23, const void* g33, const void* B,
                  const int* lx, const int* ifh2) {
  NK_INIT();
  CgState* s = as_cg(cg);
  NK_REQUIRE(s && lx && ifh2, "null argument");
  NK_REQUIRE((int64_t)s->nelv * (*lx) * (*lx) * (*lx) == s->n,
             "lx does not match the CG state size");
  NK_REQUIRE(!*ifh2 || (h2 && B), "ifh2 set but h2 or B is null");
  return ax_helm_launch<double, double, true>(
      (double*)w, (const double*)p, (const double*)dx, (const double*)h1,
      (const double*)h2, (const double*)g11, (const double*)g12,
      (const double*)g13, (const double*)g22, (const double*)g23,
      (const double*)g33, (const double*)B, *lx, s->nelv, *ifh2 != 0,
      s->part_ax, &s->sc->done);
}

int nekoamd_cg_gs_bc(void* cg, void* gs, void* w, const void* mask,
                     const int64_t* nmask) {
  NK_INIT();
  CgState* s = as_cg(cg);
  NK_REQUIRE(s && nmask, "null argument");
  int rc = gs_launch<double>(gs, (double*)w, NEKOAMD_GS_ADD, &s->sc->done);
  if (rc) return rc;
  return bc_zero_launch<double>((double*)w, (const int32_t*)mask, *nmask,
                                &s->sc->done);
}

int nekoamd_cg_alpha(void* cg) {
  NK_INIT();
  CgState* s = as_cg(cg);
  NK_REQUIRE(s, "null cg");
  hipLaunchKernelGGL(cg_alpha_k, dim3(1), dim3(kFinBlock), 0, stream(), s->sc,
                     s->part_ax, s->nelv);
  NK_LAUNCHED();
  return kOk;
}

int nekoamd_cg_update(void* cg, void* x, void* r, const void* p,
                      const void* w, const void* d, const void* mult) {
  NK_INIT();
  CgState* s = as_cg(cg);
  NK_REQUIRE(s, "null cg");
  hipLaunchKernelGGL(cg_update_k, dim3(s->grid), dim3(kRedBlock), 0, stream(),
                     s->n, (double*)x, (double*)r, (const double*)p,
                     (const double*)w, (const double*)d, (const double*)mult,
                     s->sc, s->part);
  NK_LAUNCHED();
  hipLaunchKernelGGL(cg_update_fin_k, dim3(1), dim3(kFinBlock), 0, stream(),
                     s->sc, s->part, s->grid);
  NK_LAUNCHED();
  return kOk;
}

int nekoamd_cg_iteration(void* cg, void* x, void* r, void* p, void* w,
                         const void* d, const void* mult, void* gs,
                         const void* mask, const int64_t* nmask,
                         const void* dx, const void* h1, const void* h2,
                         const void* g11, const void* g12, const void* g13,
                         const void* g22, const void* g23, const void* g33,
                         const void* B, const int* lx, const int* ifh2) {
  int rc = nekoamd_cg_pupdate(cg, p, r, d);
  if (!rc)
    rc = nekoamd_cg_ax(cg, w, p, dx, h1, h2, g11, g12, g13, g22, g23, g33, B,
                       lx, ifh2);
  if (!rc) rc = nekoamd_cg_gs_bc(cg, gs, w, mask, nmask);
  if (!rc) rc = nekoamd_cg_alpha(cg);
  if (!rc) rc = nekoamd_cg_update(cg, x, r, p, w, d, mult);
  return rc;
}

static int cg_read(CgState* s, int slot, hipEvent_t ev) {
  NK_HIP(hipMemcpyAsync(&s->host[slot], s->sc, sizeof(CgScal),
                        hipMemcpyDeviceToHost, stream()));
  NK_HIP(hipEventRecord(ev, stream()));
  return kOk;
}

int nekoamd_cg_status(void* cg, nekoamd_cg_monitor_t* mon) {
  NK_INIT();
  CgState* s = as_cg(cg);
  NK_REQUIRE(s && mon, "null argument");
  NK_HIP(hipMemcpyAsync(&s->host[0], s->sc, sizeof(CgScal),
                        hipMemcpyDeviceToHost, stream()));
  NK_HIP(hipStreamSynchronize(stream()));
  mon->iter = s->host[0].iter;
  mon->done = s->host[0].done;
  mon->res_start = s->host[0].res_start;
  mon->res_final = s->host[0].rnorm;
  return kOk;
}

int nekoamd_cg_solve(void* cg, void* x, void* r, void* p, void* w,
                     const void* f, const void* d, const void* mult,
                     const double* volume, const double* abstol,
                     const int* max_iter, const int* poll, void* gs,
                     const void* mask, const int64_t* nmask, const void* dx,
                     const void* h1, const void* h2, const void* g11,
                     const void* g12, const void* g13, const void* g22,
                     const void* g23, const void* g33, const void* B,
                     const int* lx, const int* ifh2,
                     nekoamd_cg_monitor_t* mon) {
  NK_INIT();
  CgState* s = as_cg(cg);
  NK_REQUIRE(s && poll && mon && max_iter, "null argument");
  const int batch = *poll < 1 ? 1 : *poll;
  int rc = nekoamd_cg_start(cg, x, r, f, mult, d, volume, abstol, max_iter);
  if (rc) return rc;
  hipEvent_t ev[2];
  NK_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  NK_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  // Double-buffered polling: batch b+1 is queued before the status read
  // after batch b is waited for, so the device never idles on the host.
  int cur = 0, prev = -1, queued = 0, status = kOk;
  for (;;) {
    int launched = -1;
    if (queued < *max_iter) {
      for (int q = 0; q < batch && !status; ++q)
        status = nekoamd_cg_iteration(cg, x, r, p, w, d, mult, gs, mask, nmask,
                                      dx, h1, h2, g11, g12, g13, g22, g23, g33,
                                      B, lx, ifh2);
      queued += batch;
      if (!status) status = cg_read(s, cur, ev[cur]);
      if (status) break;
      launched = cur;
    }
    if (prev >= 0) {
      hipError_t e = hipEventSynchronize(ev[prev]);
      if (e != hipSuccess) {
        status = fail_hip(__func__, e);
        break;
      }
      if (s->host[prev].done) break;
    }
    if (launched < 0) break;
    prev = launched;
    cur ^= 1;
  }
  (void)hipStreamSynchronize(stream());
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  if (status) return status;
  return nekoamd_cg_status(cg, mon);
}

}  // extern "C"
