// fcg.hip -- flexible preconditioned conjugate gradients (Polak-Ribiere
// beta) with the Jacobi preconditioner and, optionally, the element
// network's learned correction inside the iteration.
//
// It replaces ksp_prs%solve for the pressure (adjoint_pnpn.f90:734-736,
// between the projection's pre_solving and post_solving, :729-741) when a
// solver whose preconditioner varies from one iteration to the next is
// asked for.  The reference has no learned component; the method is
// flexible CG as Notay (SIAM J. Sci. Comput. 22, 2000) states it, truncated
// to one previous direction:
//   z_k = M_k(r_k),  beta_k = <z_k, r_k - r_{k-1}> / <z_{k-1}, r_{k-1}>,
//   p_k = z_k + beta_k p_{k-1},  alpha_k = <z_k, r_k> / <p_k, A p_k>,
//   x += alpha_k p_k,  r -= alpha_k A p_k.
// With r_k - r_{k-1} = -alpha_{k-1} A p_{k-1} the numerator is
// -alpha_{k-1} <z_k, A p_{k-1}>, so the previous residual need not be kept.
// For a fixed symmetric preconditioner <z_k, r_{k-1}> = 0 and the iterates
// are those of PCG.
//
// M_k(r) = D^-1 r, and every `every`-th iteration (and at the first) plus
// the learned correction gamma q:  q = mask(mult QQ^T N(r)) is the
// network's element-local prediction of A^-1 r made continuous (the
// element network of mlp.hip), gamma = <q, r> / <q, A q> its Galerkin
// weight (0 when <q, A q> <= 0): the coarse correction of the learned aid.
//
// All scalars stay on the device; every kernel returns at once when the
// solve is done; the host reads the state once per batch of `poll`
// iterations (double-buffered, as the CG driver does).
#include "internal.h"
#include "nekoamd.h"
#include "reduce.h"

namespace nekoamd {

struct FcgScal {
  double rz, rz_new, zw, alpha, beta, pap, rnorm, res_start, norm_fac, abstol;
  double qr, qaq, gamma;
  int iter, max_iter, done, pad;
};

struct FcgState {
  int64_t n = 0;
  int nelv = 0;
  int grid = 1;
  FcgScal* sc = nullptr;     // device
  FcgScal* host = nullptr;   // pinned, 2 slots
  double *r = nullptr, *z = nullptr, *p = nullptr, *w = nullptr;
  double *q = nullptr, *aq = nullptr, *dm = nullptr;
  double* part = nullptr;    // 2 * kRedGrid
};

// x = 0, r = f; partials of <r, r>.
__global__ void __launch_bounds__(kRedBlock)
    fcg_start_k(int64_t n, double* __restrict__ x, double* __restrict__ r,
                const double* __restrict__ f, const double* __restrict__ mult,
                double* __restrict__ part) {
  __shared__ double sh[kRedBlock / 64];
  double a = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRedBlock) {
    const double fi = f[i];
    x[i] = 0.0;
    r[i] = fi;
    a += mult[i] * fi * fi;
  }
  a = block_sum<kRedBlock>(a, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}

__global__ void __launch_bounds__(kFinBlock)
    fcg_start_fin_k(FcgScal* __restrict__ sc, const double* __restrict__ part,
                    int np, double volume, double abstol, int max_iter) {
  __shared__ double sh[kFinBlock / 64];
  const double rr = sum_partials<kFinBlock>(part, np, sh);
  if (threadIdx.x == 0) {
    const double nf = 1.0 / sqrt(volume);
    const double rn = sqrt(rr) * nf;
    FcgScal s{};
    s.norm_fac = nf;
    s.abstol = abstol;
    s.rnorm = rn;
    s.res_start = rn;
    s.max_iter = max_iter;
    s.done = (rn == 0.0 || max_iter <= 0) ? 1 : 0;
    *sc = s;
  }
}

// Learned correction, Galerkin weight: partials of <q, r> and <q, A q>.
__global__ void __launch_bounds__(kRedBlock)
    fcg_qdots_k(int64_t n, const double* __restrict__ q,
                const double* __restrict__ r, const double* __restrict__ aq,
                const double* __restrict__ mult,
                const FcgScal* __restrict__ sc, double* __restrict__ part) {
  __shared__ double sh[kRedBlock / 64];
  if (sc->done) return;
  double a = 0.0, b = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRedBlock) {
    const double m = mult[i] * q[i];
    a += m * r[i];
    b += m * aq[i];
  }
  a = block_sum<kRedBlock>(a, sh);
  b = block_sum<kRedBlock>(b, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[kRedGrid + blockIdx.x] = b;
  }
}

__global__ void __launch_bounds__(kFinBlock)
    fcg_gamma_k(FcgScal* __restrict__ sc, const double* __restrict__ part,
                int np) {
  if (sc->done) return;
  __shared__ double sh[kFinBlock / 64];
  const double qr = sum_partials<kFinBlock>(part, np, sh);
  const double qaq = sum_partials<kFinBlock>(part + kRedGrid, np, sh);
  if (threadIdx.x == 0) {
    const double g = qaq > 0.0 ? qr / qaq : 0.0;
    sc->qr = qr;
    sc->qaq = qaq;
    sc->gamma = isfinite(g) ? g : 0.0;
  }
}

// z = D^-1 r (+ gamma q when `learned`); partials of <z, r> and of
// <z, w> (w = A p_{k-1}, still in its buffer).
__global__ void __launch_bounds__(kRedBlock)
    fcg_prec_k(int64_t n, double* __restrict__ z, const double* __restrict__ r,
               const double* __restrict__ dm, const double* __restrict__ q,
               const double* __restrict__ w, const double* __restrict__ mult,
               const FcgScal* __restrict__ sc, int learned,
               double* __restrict__ part) {
  __shared__ double sh[kRedBlock / 64];
  if (sc->done) return;
  const double g = learned ? sc->gamma : 0.0;
  const bool first = sc->iter == 0;
  double a = 0.0, b = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRedBlock) {
    double zi = dm[i] * r[i];
    if (learned) zi += g * q[i];
    z[i] = zi;
    const double m = mult[i] * zi;
    a += m * r[i];
    if (!first) b += m * w[i];
  }
  a = block_sum<kRedBlock>(a, sh);
  b = block_sum<kRedBlock>(b, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[kRedGrid + blockIdx.x] = b;
  }
}

// beta = -alpha_{k-1} <z_k, A p_{k-1}> / <z_{k-1}, r_{k-1}> (0 at the first
// iteration); rz = <z_k, r_k>.
__global__ void __launch_bounds__(kFinBlock)
    fcg_beta_k(FcgScal* __restrict__ sc, const double* __restrict__ part,
               int np) {
  if (sc->done) return;
  __shared__ double sh[kFinBlock / 64];
  const double zr = sum_partials<kFinBlock>(part, np, sh);
  const double zw = sum_partials<kFinBlock>(part + kRedGrid, np, sh);
  if (threadIdx.x == 0) {
    const bool first = sc->iter == 0;
    sc->zw = zw;
    sc->beta = first ? 0.0 : -sc->alpha * zw / sc->rz;
    sc->rz = zr;
  }
}

// p = z + beta p (p = z at the first iteration: p holds no data yet).
__global__ void __launch_bounds__(256)
    fcg_dir_k(int64_t n, double* __restrict__ p, const double* __restrict__ z,
              const FcgScal* __restrict__ sc) {
  if (sc->done) return;
  const bool first = sc->iter == 0;
  const double beta = sc->beta;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256)
    p[i] = first ? z[i] : z[i] + beta * p[i];
}

// partials of <p, w>.
__global__ void __launch_bounds__(kRedBlock)
    fcg_pw_k(int64_t n, const double* __restrict__ p,
             const double* __restrict__ w, const double* __restrict__ mult,
             const FcgScal* __restrict__ sc, double* __restrict__ part) {
  __shared__ double sh[kRedBlock / 64];
  if (sc->done) return;
  double a = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRedBlock)
    a += mult[i] * p[i] * w[i];
  a = block_sum<kRedBlock>(a, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}

__global__ void __launch_bounds__(kFinBlock)
    fcg_alpha_k(FcgScal* __restrict__ sc, const double* __restrict__ part,
                int np) {
  if (sc->done) return;
  __shared__ double sh[kFinBlock / 64];
  const double pap = sum_partials<kFinBlock>(part, np, sh);
  if (threadIdx.x == 0) {
    sc->pap = pap;
    sc->alpha = sc->rz / pap;
  }
}

// x += alpha p, r -= alpha w; partials of <r, r>.
__global__ void __launch_bounds__(kRedBlock)
    fcg_update_k(int64_t n, double* __restrict__ x, double* __restrict__ r,
                 const double* __restrict__ p, const double* __restrict__ w,
                 const double* __restrict__ mult,
                 const FcgScal* __restrict__ sc, double* __restrict__ part) {
  __shared__ double sh[kRedBlock / 64];
  if (sc->done) return;
  const double alpha = sc->alpha;
  double a = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRedBlock) {
    x[i] = x[i] + alpha * p[i];
    const double ri = r[i] - alpha * w[i];
    r[i] = ri;
    a += mult[i] * ri * ri;
  }
  a = block_sum<kRedBlock>(a, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = a;
}

__global__ void __launch_bounds__(kFinBlock)
    fcg_update_fin_k(FcgScal* __restrict__ sc, const double* __restrict__ part,
                     int np) {
  if (sc->done) return;
  __shared__ double sh[kFinBlock / 64];
  const double rr = sum_partials<kFinBlock>(part, np, sh);
  if (threadIdx.x == 0) {
    const int it = sc->iter + 1;
    const double rn = sqrt(rr) * sc->norm_fac;
    sc->iter = it;
    sc->rnorm = rn;
    if (rn < sc->abstol || it >= sc->max_iter || !isfinite(rn)) sc->done = 1;
  }
}

__global__ void __launch_bounds__(256)
    fcg_mask_copy_k(int64_t n, double* __restrict__ dm,
                    const double* __restrict__ d) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256)
    dm[i] = d[i];
}

static FcgState* as_fcg(void* p) { return (FcgState*)p; }

}  // namespace nekoamd

using namespace nekoamd;

extern "C" {

int nekoamd_fcg_create(void** st, const int64_t* n, const int* nelv) {
  NK_INIT();
  NK_REQUIRE(st && n && nelv && *n >= 0 && *nelv >= 0, "bad arguments");
  FcgState* s = new FcgState();
  s->n = *n;
  s->nelv = *nelv;
  s->grid = nk_grid(*n, kRedBlock, kRedGrid);
  const size_t nb = sizeof(double) * (size_t)(*n > 0 ? *n : 1);
  hipError_t e = hipMalloc((void**)&s->sc, sizeof(FcgScal));
  double** vecs[7] = {&s->r, &s->z, &s->p, &s->w, &s->q, &s->aq, &s->dm};
  for (double** v : vecs)
    if (e == hipSuccess) e = hipMalloc((void**)v, nb);
  if (e == hipSuccess)
    e = hipMalloc((void**)&s->part, sizeof(double) * 2 * kRedGrid);
  if (e == hipSuccess)
    e = hipHostMalloc((void**)&s->host, 2 * sizeof(FcgScal),
                      hipHostMallocDefault);
  if (e == hipSuccess) {
    FcgScal init{};
    init.done = 1;
    e = hipMemcpy(s->sc, &init, sizeof(FcgScal), hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) {
    (void)nekoamd_fcg_free(s);
    return fail_hip(__func__, e);
  }
  *st = s;
  return kOk;
}

int nekoamd_fcg_free(void* st) {
  FcgState* s = as_fcg(st);
  if (!s) return kOk;
  if (initialised()) (void)hipStreamSynchronize(stream());
  for (double* v : {s->r, s->z, s->p, s->w, s->q, s->aq, s->dm, s->part})
    if (v) (void)hipFree(v);
  if (s->sc) (void)hipFree(s->sc);
  if (s->host) (void)hipHostFree(s->host);
  delete s;
  return kOk;
}

int nekoamd_fcg_solve(void* st, void* x, const void* f, const void* d,
                      const void* mult, const double* volume,
                      const double* abstol, const int* max_iter,
                      const int* poll, void* gs, const void* mask,
                      const int64_t* nmask, const void* dx, const void* h1,
                      const void* h2, const void* g11, const void* g12,
                      const void* g13, const void* g22, const void* g23,
                      const void* g33, const void* B, const int* lx,
                      const int* ifh2, void* mlp, const int* every,
                      nekoamd_cg_monitor_t* mon) {
  NK_INIT();
  FcgState* s = as_fcg(st);
  NK_REQUIRE(s && volume && abstol && max_iter && poll && nmask && lx &&
                 ifh2 && every && mon,
             "null argument");
  NK_REQUIRE(*volume > 0.0, "volume must be positive");
  NK_REQUIRE(*nmask >= 0 && (*nmask == 0 || mask), "bad mask");
  NK_REQUIRE(*lx >= 2 && *lx <= NEKOAMD_LX_MAX, "lx must be in [2, 10]");
  NK_REQUIRE((int64_t)s->nelv * (*lx) * (*lx) * (*lx) == s->n,
             "lx does not match the FCG state size");
  NK_REQUIRE(s->n == 0 || (x && f && d && mult && gs && dx && h1 && g11 &&
                           g12 && g13 && g22 && g23 && g33),
             "null field");
  NK_REQUIRE(!*ifh2 || (h2 && B), "ifh2 set but h2 or B is null");
  NK_REQUIRE(!mlp || *lx == 8, "the learned correction needs lx = 8");
  NK_REQUIRE(*every >= 1, "every must be positive");
  const int64_t n = s->n;
  const int ge = nk_grid(n, 256);
  const int G = s->grid;
  const int32_t* M = (const int32_t*)mask;
  double* X = (double*)x;
  const double* F = (const double*)f;
  const double* MU = (const double*)mult;
  const int* done = &s->sc->done;
  auto op = [&](double* out, const double* in) -> int {
    int rc = ax_helm_launch<double, double>(
        out, in, (const double*)dx, (const double*)h1, (const double*)h2,
        (const double*)g11, (const double*)g12, (const double*)g13,
        (const double*)g22, (const double*)g23, (const double*)g33,
        (const double*)B, *lx, s->nelv, *ifh2 != 0, done);
    if (!rc) rc = gs_launch<double>(gs, out, NEKOAMD_GS_ADD, done);
    if (!rc) rc = bc_zero_launch<double>(out, M, *nmask, done);
    return rc;
  };
  hipLaunchKernelGGL(fcg_mask_copy_k, dim3(ge), dim3(256), 0, stream(), n,
                     s->dm, (const double*)d);
  NK_LAUNCHED();
  int rc = bc_zero_launch<double>(s->dm, M, *nmask, nullptr);
  if (rc) return rc;
  hipLaunchKernelGGL(fcg_start_k, dim3(G), dim3(kRedBlock), 0, stream(), n, X,
                     s->r, F, MU, s->part);
  NK_LAUNCHED();
  hipLaunchKernelGGL(fcg_start_fin_k, dim3(1), dim3(kFinBlock), 0, stream(),
                     s->sc, s->part, G, *volume, *abstol, *max_iter);
  NK_LAUNCHED();
  int k = 0;  // iterations queued
  auto iteration = [&]() -> int {
    const int learned = mlp != nullptr && k % *every == 0;
    if (learned) {
      int r2 = mlp_direction(mlp, n, s->r, s->q, MU, gs, M, *nmask);
      if (!r2) r2 = op(s->aq, s->q);
      if (r2) return r2;
      hipLaunchKernelGGL(fcg_qdots_k, dim3(G), dim3(kRedBlock), 0, stream(),
                         n, s->q, s->r, s->aq, MU, s->sc, s->part);
      NK_LAUNCHED();
      hipLaunchKernelGGL(fcg_gamma_k, dim3(1), dim3(kFinBlock), 0, stream(),
                         s->sc, s->part, G);
      NK_LAUNCHED();
    }
    hipLaunchKernelGGL(fcg_prec_k, dim3(G), dim3(kRedBlock), 0, stream(), n,
                       s->z, s->r, s->dm, s->q, s->w, MU, s->sc, learned,
                       s->part);
    NK_LAUNCHED();
    hipLaunchKernelGGL(fcg_beta_k, dim3(1), dim3(kFinBlock), 0, stream(),
                       s->sc, s->part, G);
    NK_LAUNCHED();
    hipLaunchKernelGGL(fcg_dir_k, dim3(ge), dim3(256), 0, stream(), n, s->p,
                       s->z, s->sc);
    NK_LAUNCHED();
    int r3 = op(s->w, s->p);
    if (r3) return r3;
    hipLaunchKernelGGL(fcg_pw_k, dim3(G), dim3(kRedBlock), 0, stream(), n,
                       s->p, s->w, MU, s->sc, s->part);
    NK_LAUNCHED();
    hipLaunchKernelGGL(fcg_alpha_k, dim3(1), dim3(kFinBlock), 0, stream(),
                       s->sc, s->part, G);
    NK_LAUNCHED();
    hipLaunchKernelGGL(fcg_update_k, dim3(G), dim3(kRedBlock), 0, stream(), n,
                       X, s->r, s->p, s->w, MU, s->sc, s->part);
    NK_LAUNCHED();
    hipLaunchKernelGGL(fcg_update_fin_k, dim3(1), dim3(kFinBlock), 0,
                       stream(), s->sc, s->part, G);
    NK_LAUNCHED();
    ++k;
    return kOk;
  };
  // batches of `poll` iterations; the status read after batch b is waited
  // for once batch b+1 is queued
  hipEvent_t ev[2];
  NK_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  hipError_t e1 = hipEventCreateWithFlags(&ev[1], hipEventDisableTiming);
  if (e1 != hipSuccess) {
    (void)hipEventDestroy(ev[0]);
    return fail_hip(__func__, e1);
  }
  const int batch = *poll < 1 ? 1 : *poll;
  int cur = 0, pending = -1, status = kOk;
  while (status == kOk) {
    int launched = -1;
    if (k < *max_iter) {
      for (int q = 0; q < batch && status == kOk; ++q) status = iteration();
      if (status) break;
      hipError_t e = hipMemcpyAsync(&s->host[cur], s->sc, sizeof(FcgScal),
                                    hipMemcpyDeviceToHost, stream());
      if (e == hipSuccess) e = hipEventRecord(ev[cur], stream());
      if (e != hipSuccess) {
        status = fail_hip(__func__, e);
        break;
      }
      launched = cur;
      cur ^= 1;
    }
    if (pending >= 0) {
      const hipError_t e = hipEventSynchronize(ev[pending]);
      if (e != hipSuccess) {
        status = fail_hip(__func__, e);
        break;
      }
      if (s->host[pending].done) break;
    }
    if (launched < 0) break;
    pending = launched;
  }
  (void)hipStreamSynchronize(stream());
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  if (status) return status;
  FcgScal h;
  NK_HIP(hipMemcpy(&h, s->sc, sizeof(FcgScal), hipMemcpyDeviceToHost));
  mon->iter = h.iter;
  mon->done = h.done;
  mon->res_start = h.res_start;
  mon->res_final = h.rnorm;
  mon->nupdate = 0;
  mon->pad = 0;
  return kOk;
}

}  // extern "C"
