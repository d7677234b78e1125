// gmres.hip -- restarted flexible GMRES, Neko's gmres_t: the 'gmres' of
// krylov_solver_factory that the shipped cases use for the pressure
// (examples/rugby_ball/rugby_ball.case:46-52, with 'hsmg'), called where the
// reference solves (adjoint_pnpn.f90:734-736).  Neko is not vendored; the
// iteration is restated from its algorithm:
//   x = 0, r = f;  outer: gam_1 = ||r||_mult, v_1 = r / gam_1
//   inner j = 1..m:  z_j = M^-1 v_j;  w = mask(QQ^T A_L z_j)
//                    h_ij = <w, v_i>_mult (i <= j, classical Gram-Schmidt),
//                    w -= sum_i h_ij v_i;  Givens rotations of column j;
//                    alpha = ||w||_mult;  new rotation (c_j, s_j);
//                    gam_{j+1} = -s_j gam_j, gam_j = c_j gam_j;
//                    rnorm = |gam_{j+1}| / sqrt(volume): stop below abstol;
//                    v_{j+1} = w / alpha
//   x += sum_j y_j z_j with H y = gam (back substitution); restart with
//   r = f - mask(QQ^T A_L x).
// Flexible: the preconditioned vectors z_j are kept, so M may change from
// one application to the next (the hybrid Schwarz multigrid does).
//
// Per inner iteration the device runs the preconditioner, the operator
// (ax_helm + gs + mask), one batched kernel for all j dot products against
// the basis (w read once per basis vector, partials per block) and one fused
// kernel for the orthogonalisation and the partials of ||w||^2; the j + 2
// scalars come to the host once (one sync per iteration) for the rotations,
// which cost nothing.  The basis is stored unnormalised: slot i holds
// V_i = sv_i v_i (the orthogonalised vector as the fused kernel wrote it,
// or the residual at a restart), so no pass scales it.  The preconditioner
// and the operator are applied to V_j (z_j = Z_j / sv_j), the dot products
// and coefficients carry the scales (h_ij = <A Z_j, V_i> / (sv_i sv_j); the
// kernel removes (h'_i / sv_i^2) V_i), and the update adds (y_j / sv_j) Z_j.
#include <cmath>
#include <type_traits>
#include <vector>

#include "internal.h"
#include "mdot.h"
#include "nekoamd.h"
#include "reduce.h"

namespace nekoamd {

int hsmg_apply(void* hsmg, double* z, const double* r);

namespace {

constexpr int kGmBlock = 256;

struct GmresState {
  int64_t n = 0;
  int m = 0, G = 0;
  double* V = nullptr;    // m + 1 unnormalised basis vectors of n
  double* Z = nullptr;    // m preconditioned vectors of n
  double* w = nullptr;
  double* part = nullptr;  // (m + 1) * G partial sums
  double* sc = nullptr;    // m + 1 device scalars: the dots
  double* nv = nullptr;    // m + 1 device scalars: <V_i, V_i>_mult
  // pinned host memory: [0, m + 2) the scalars the device reports (dots and
  // the new norm, written by gm_report_k; the multiplicity check's count)
  double* host = nullptr;
  std::vector<double> H, c, s, gam, y, sv;
  // the multiplicity weights as one-byte counts (mult = 1 / count exactly),
  // built on the first solve with a given mult array (keyed by its address:
  // the multiplicity of a mesh is fixed once built); cnt_ok when every point
  // passed the check, else the passes read the FP64 weights
  uint8_t* cnt8 = nullptr;
  int* bad = nullptr;
  const void* cnt_key = nullptr;
  bool cnt_ok = false;
  nekoamd_cg_monitor_t mon{};
};

// part[i*G + b] = block b's share of <a, V_i>_mult, i = blockIdx.y
__global__ void __launch_bounds__(kGmBlock)
    gm_mdot_k(int64_t n, const double* __restrict__ a,
              const double* __restrict__ V, const double* __restrict__ mult,
              double* __restrict__ part) {
  __shared__ double sh[kGmBlock / 64];
  const double* v = V + (int64_t)blockIdx.y * n;
  double s = 0.0;
  for (int64_t p = (int64_t)blockIdx.x * kGmBlock + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * kGmBlock)
    s += a[p] * v[p] * mult[p];
  s = block_sum<kGmBlock>(s, sh);
  if (threadIdx.x == 0) part[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = s;
}

__global__ void __launch_bounds__(kFinBlock)
    gm_fin_k(const double* __restrict__ part, int G, double* __restrict__ out) {
  __shared__ double sh[kFinBlock / 64];
  const double s = sum_partials<kFinBlock>(part + (int64_t)blockIdx.x * G, G, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// nv = the sum of the G partials (the new basis vector's norm squared), and
// the iteration's report written straight to pinned host memory: hout[0..k)
// = the dots dots[0..k), hout[k] = nv.  The host reads it once the kernel
// has completed (its writes are visible then), with no copy kernels.
__global__ void __launch_bounds__(kFinBlock)
    gm_report_k(const double* __restrict__ part, int G, double* __restrict__ nv,
                const double* __restrict__ dots, int k, double* __restrict__ hout) {
  __shared__ double sh[kFinBlock / 64];
  const double s = sum_partials<kFinBlock>(part, G, sh);
  if (threadIdx.x < k) hout[threadIdx.x] = dots[threadIdx.x];
  if (threadIdx.x == 0) {
    *nv = s;
    hout[k] = s;
  }
}

// out = w - sum_{i<k} (h_i / nv_i) V_i (nv_i = <V_i, V_i>_mult), partials
// of <out, out>_mult.  P2: point pairs with 16-byte loads (n even, aligned
// bases), the basis streamed nontemporally.
template <int KM, bool P2, typename TM>
__global__ void __launch_bounds__(kGmBlock)
    gm_orth_k(int64_t n, int k, const double* __restrict__ h,
              const double* __restrict__ nv, const double* __restrict__ V,
              const double* __restrict__ w, double* __restrict__ out,
              const TM* __restrict__ mult, double* __restrict__ part) {
  __shared__ double sh[kGmBlock / 64];
  __shared__ double sk[64];
  if (threadIdx.x < k) sk[threadIdx.x] = h[threadIdx.x] / nv[threadIdx.x];
  __syncthreads();
  double s = 0.0;
  const int64_t nq = P2 ? n / 2 : n;
  for (int64_t q = (int64_t)blockIdx.x * kGmBlock + threadIdx.x; q < nq;
       q += (int64_t)gridDim.x * kGmBlock) {
    double2 c = make_double2(0.0, 0.0), unused = c;
    comb_point<KM, P2, false>(sk, V, nullptr, n, k, q, c, unused);
    if constexpr (P2) {
      const double2 wo = ((const double2*)w)[q], mv = mult_pair(mult, q);
      const double2 a = make_double2(wo.x - c.x, wo.y - c.y);
      ((double2*)out)[q] = a;
      s += a.x * a.x * mv.x + a.y * a.y * mv.y;
    } else {
      const double a = w[q] - c.x;
      out[q] = a;
      s += a * a * mult_at(mult, q);
    }
  }
  s = block_sum<kGmBlock>(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

template <int KM, typename TM>
void orth_launch(int G, int64_t n, int k, const double* h, const double* nv,
                 const double* V, const double* w, double* out,
                 const TM* mult, double* part, bool p2) {
  if (p2)
    hipLaunchKernelGGL((gm_orth_k<KM, true, TM>), dim3(G), dim3(kGmBlock), 0,
                       stream(), n, k, h, nv, V, w, out, mult, part);
  else
    hipLaunchKernelGGL((gm_orth_k<KM, false, TM>), dim3(G), dim3(kGmBlock), 0,
                       stream(), n, k, h, nv, V, w, out, mult, part);
}

// out = w - sum_{i<k} (h_i / nv_i) V_i and the partials of ||out||^2_mult,
// one pass.
template <typename TM>
void orth(int G, int64_t n, int k, const double* h, const double* nv,
          const double* V, const double* w, double* out, const TM* mult,
          double* part) {
  const bool p2 = pairs_ok(n, V, w, out, sizeof(TM) == 8 ? (const void*)mult : nullptr) &&
                  ((uintptr_t)mult & (2 * sizeof(TM) - 1)) == 0;
  if (k <= 8)
    orth_launch<8>(G, n, k, h, nv, V, w, out, mult, part, p2);
  else if (k <= 16)
    orth_launch<16>(G, n, k, h, nv, V, w, out, mult, part, p2);
  else if (k <= 32)
    orth_launch<32>(G, n, k, h, nv, V, w, out, mult, part, p2);
  else
    orth_launch<64>(G, n, k, h, nv, V, w, out, mult, part, p2);
}

// The update's coefficients, passed by value (no upload).
struct GmCoef {
  double c[NEKOAMD_GMRES_MAX_M];
};

// x += sum_{i<k} y_i Z_i (SET: x = the sum, x not read), with the point
// pairs and nontemporal basis loads of the orthogonalisation.
template <int KM, bool P2, bool SET>
__global__ void __launch_bounds__(kGmBlock)
    gm_update_k(int64_t n, int k, GmCoef y, const double* __restrict__ Z,
                double* __restrict__ x) {
  __shared__ double sk[64];
  if (threadIdx.x < k) sk[threadIdx.x] = y.c[threadIdx.x];
  __syncthreads();
  const int64_t nq = P2 ? n / 2 : n;
  for (int64_t q = (int64_t)blockIdx.x * kGmBlock + threadIdx.x; q < nq;
       q += (int64_t)gridDim.x * kGmBlock) {
    // the sum starts from x (0 when SET) and adds y_i Z_i in order, as a
    // sequential x += y_i Z_i would
    double2 c = make_double2(0.0, 0.0), unused = c;
    if constexpr (!SET) c = P2 ? ((const double2*)x)[q] : make_double2(x[q], 0.0);
    comb_point<KM, P2, false>(sk, Z, nullptr, n, k, q, c, unused);
    if constexpr (P2)
      ((double2*)x)[q] = c;
    else
      x[q] = c.x;
  }
}

template <int KM, bool SET>
void update_launch(int64_t n, int k, const GmCoef& y, const double* Z,
                   double* x, bool p2) {
  const int G = nk_grid(n, kGmBlock, kMultiGrid);
  if (p2)
    hipLaunchKernelGGL((gm_update_k<KM, true, SET>), dim3(G), dim3(kGmBlock), 0,
                       stream(), n, k, y, Z, x);
  else
    hipLaunchKernelGGL((gm_update_k<KM, false, SET>), dim3(G), dim3(kGmBlock),
                       0, stream(), n, k, y, Z, x);
}

// x (+)= sum_{i<k} y_i Z_i; set: x is not read (the first cycle's update)
void update(int64_t n, int k, const GmCoef& y, const double* Z, double* x,
            bool set) {
  const bool p2 = pairs_ok(n, Z, x);
  auto go = [&](auto km) {
    constexpr int K = decltype(km)::value;
    if (set)
      update_launch<K, true>(n, k, y, Z, x, p2);
    else
      update_launch<K, false>(n, k, y, Z, x, p2);
  };
  if (k <= 8)
    go(std::integral_constant<int, 8>{});
  else if (k <= 16)
    go(std::integral_constant<int, 16>{});
  else if (k <= 32)
    go(std::integral_constant<int, 32>{});
  else
    go(std::integral_constant<int, 64>{});
}

// out = s a  (s = 0 zeroes); out may be a
__global__ void __launch_bounds__(kGmBlock)
    gm_scale_k(int64_t n, double* out, const double* a, double s) {
  for (int64_t p = (int64_t)blockIdx.x * kGmBlock + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * kGmBlock)
    out[p] = s == 0.0 ? 0.0 : s * a[p];
}

// out = a b
__global__ void __launch_bounds__(kGmBlock)
    gm_mul_k(int64_t n, double* __restrict__ out, const double* __restrict__ a,
             const double* __restrict__ b) {
  for (int64_t p = (int64_t)blockIdx.x * kGmBlock + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * kGmBlock)
    out[p] = a[p] * b[p];
}

// r = f - w
__global__ void __launch_bounds__(kGmBlock)
    gm_sub_k(int64_t n, double* __restrict__ r, const double* __restrict__ f,
             const double* __restrict__ w) {
  for (int64_t p = (int64_t)blockIdx.x * kGmBlock + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * kGmBlock)
    r[p] = f[p] - w[p];
}

struct Op {
  void* gs;
  const int32_t* mask;
  int64_t nmask;
  const double *dx, *h1, *h2, *g11, *g12, *g13, *g22, *g23, *g33, *B;
  int lx, nelv;
  bool ifh2;
};

int apply_op(const Op& op, double* w, const double* u) {
  int rc = ax_helm_launch<double, double>(w, u, op.dx, op.h1, op.h2, op.g11,
                                          op.g12, op.g13, op.g22, op.g23,
                                          op.g33, op.B, op.lx, op.nelv,
                                          op.ifh2, nullptr);
  if (rc != kOk) return rc;
  rc = gs_launch<double>(op.gs, w, 1, nullptr);
  if (rc != kOk) return rc;
  return bc_zero_launch<double>(w, op.mask, op.nmask, nullptr);
}

// *dev = <a, a>_mult on the device and in s->host[0] (synchronises)
int mdot(GmresState* s, const double* a, const double* mult, double* dev) {
  hipLaunchKernelGGL(gm_mdot_k, dim3(s->G, 1), dim3(kGmBlock), 0, stream(),
                     s->n, a, a, mult, s->part);
  hipLaunchKernelGGL(gm_report_k, dim3(1), dim3(kFinBlock), 0, stream(), s->part,
                     s->G, dev, (const double*)nullptr, 0, s->host);
  NK_LAUNCHED();
  NK_HIP(hipStreamSynchronize(stream()));
  return kOk;
}

int precond(int kind, void* pc, int64_t n, double* z, const double* v) {
  const int g = nk_grid(n, kGmBlock);
  switch (kind) {
    case NEKOAMD_PC_IDENT:
      NK_HIP(hipMemcpyAsync(z, v, sizeof(double) * n, hipMemcpyDeviceToDevice,
                            stream()));
      return kOk;
    case NEKOAMD_PC_JACOBI:
      hipLaunchKernelGGL(gm_mul_k, dim3(g), dim3(kGmBlock), 0, stream(), n, z,
                         (const double*)pc, v);
      NK_LAUNCHED();
      return kOk;
    case NEKOAMD_PC_HSMG:
      return hsmg_apply(pc, z, v);
    default:
      return fail_arg("nekoamd_gmres_solve", "unknown preconditioner kind");
  }
}

}  // namespace
}  // namespace nekoamd

using namespace nekoamd;

extern "C" {

int nekoamd_gmres_create(void** st, const int64_t* n, const int* m) {
  NK_INIT();
  NK_REQUIRE(st && n && m, "null argument");
  NK_REQUIRE(*n >= 0, "n < 0");
  NK_REQUIRE(*m >= 1 && *m <= NEKOAMD_GMRES_MAX_M, "restart length out of range");
  auto* s = new GmresState;
  s->n = *n;
  s->m = *m;
  s->G = nk_grid(*n, kGmBlock, kMultiGrid);
  const size_t nb = sizeof(double) * (size_t)std::max<int64_t>(*n, 1);
  hipError_t e = hipMalloc(&s->V, nb * (s->m + 1));
  if (e == hipSuccess) e = hipMalloc(&s->Z, nb * s->m);
  if (e == hipSuccess) e = hipMalloc(&s->w, nb);
  if (e == hipSuccess)
    e = hipMalloc(&s->part, sizeof(double) * (size_t)(s->m + 1) * s->G);
  if (e == hipSuccess) e = hipMalloc(&s->sc, sizeof(double) * (s->m + 1));
  if (e == hipSuccess) e = hipMalloc(&s->nv, sizeof(double) * (s->m + 1));
  if (e == hipSuccess)
    e = hipHostMalloc((void**)&s->host, sizeof(double) * (s->m + 2),
                      hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc(&s->bad, sizeof(int));
  if (e != hipSuccess) {
    (void)hipHostFree(s->host);
    (void)hipFree(s->bad);
    (void)hipFree(s->V);
    (void)hipFree(s->Z);
    (void)hipFree(s->w);
    (void)hipFree(s->part);
    (void)hipFree(s->sc);
    (void)hipFree(s->nv);
    delete s;
    return fail_hip(__func__, e);
  }
  const int m1 = s->m + 1;
  s->H.assign((size_t)m1 * m1, 0.0);
  s->c.assign(m1, 0.0);
  s->s.assign(m1, 0.0);
  s->gam.assign(m1 + 1, 0.0);
  s->y.assign(m1, 0.0);
  s->sv.assign(m1, 0.0);
  *st = s;
  return kOk;
}

int nekoamd_gmres_free(void* st) {
  NK_INIT();
  auto* s = static_cast<GmresState*>(st);
  if (!s) return kOk;
  NK_HIP(hipStreamSynchronize(stream()));
  NK_HIP(hipFree(s->V));
  NK_HIP(hipFree(s->Z));
  NK_HIP(hipFree(s->w));
  NK_HIP(hipFree(s->part));
  NK_HIP(hipFree(s->sc));
  NK_HIP(hipFree(s->nv));
  NK_HIP(hipHostFree(s->host));
  NK_HIP(hipFree(s->bad));
  NK_HIP(hipFree(s->cnt8));
  delete s;
  return kOk;
}

int nekoamd_gmres_status(void* st, nekoamd_cg_monitor_t* mon) {
  auto* s = static_cast<GmresState*>(st);
  NK_REQUIRE(s && mon, "null argument");
  *mon = s->mon;
  return kOk;
}

int nekoamd_gmres_solve(void* st, void* x_, const void* f_, const int* pc_kind,
                        void* pc, const void* mult_, const double* volume,
                        const double* abstol, const int* max_iter, void* gs,
                        const void* mask, const int64_t* nmask, const void* dx,
                        const void* h1, const void* h2, const void* g11,
                        const void* g12, const void* g13, const void* g22,
                        const void* g23, const void* g33, const void* B,
                        const int* lx, const int* ifh2,
                        nekoamd_cg_monitor_t* mon) {
  NK_INIT();
  auto* s = static_cast<GmresState*>(st);
  NK_REQUIRE(s && pc_kind && volume && abstol && max_iter && nmask && lx && ifh2,
             "null argument");
  NK_REQUIRE(*volume > 0.0, "volume <= 0");
  NK_REQUIRE(*pc_kind == NEKOAMD_PC_IDENT || pc != nullptr, "null preconditioner");
  const int64_t n = s->n;
  const int L = *lx;
  NK_REQUIRE(L >= 2 && n % ((int64_t)L * L * L) == 0, "n is not nelv lx^3");
  auto* x = (double*)x_;
  auto* f = (const double*)f_;
  auto* mult = (const double*)mult_;
  s->mon = nekoamd_cg_monitor_t{};
  if (n == 0) {
    if (mon) *mon = s->mon;
    return kOk;
  }
  NK_REQUIRE(x && f && mult && gs && dx && h1 && g11 && g22 && g33 && B,
             "null field");
  NK_REQUIRE(*nmask == 0 || mask, "null mask");
  Op op{gs, (const int32_t*)mask, *nmask, (const double*)dx,
        (const double*)h1, (const double*)h2, (const double*)g11,
        (const double*)g12, (const double*)g13, (const double*)g22,
        (const double*)g23, (const double*)g33, (const double*)B, L,
        (int)(n / ((int64_t)L * L * L)), *ifh2 != 0};
  const int g = nk_grid(n, kGmBlock);
  const int m = s->m;
  const double norm_fac = 1.0 / std::sqrt(*volume);
  auto H = [&](int i, int j) -> double& { return s->H[(size_t)i * (m + 1) + j]; };

  if (mult != s->cnt_key) {
    if (!s->cnt8) NK_HIP(hipMalloc((void**)&s->cnt8, (size_t)n));
    if (const int rc0 = mult_counts(n, mult, s->cnt8, s->bad, (int*)s->host,
                                    &s->cnt_ok))
      return rc0;
    s->cnt_key = mult;
  }
  // x is written by the first cycle's update (x = the sum), or zeroed at the
  // end when no update ran
  bool xset = false;
  int iter = 0, rc = kOk;
  bool conv = false;
  double rnorm = 0.0;
  while (!conv && iter < *max_iter) {
    // V_0 = r = f (first cycle) or f - A x, sv_0 = ||r||
    if (iter == 0) {
      NK_HIP(hipMemcpyAsync(s->V, f, sizeof(double) * n,
                            hipMemcpyDeviceToDevice, stream()));
    } else {
      if ((rc = apply_op(op, s->w, x)) != kOk) return rc;
      hipLaunchKernelGGL(gm_sub_k, dim3(g), dim3(kGmBlock), 0, stream(), n,
                         s->V, f, s->w);
      NK_LAUNCHED();
    }
    if ((rc = mdot(s, s->V, mult, s->nv)) != kOk) return rc;
    s->gam.assign(m + 2, 0.0);
    s->gam[0] = std::sqrt(s->host[0]);
    s->sv[0] = s->gam[0];
    if (iter == 0) s->mon.res_start = s->gam[0] * norm_fac;
    if (s->gam[0] == 0.0) {
      rnorm = 0.0;
      conv = true;
      break;
    }
    if (!std::isfinite(s->gam[0])) {  // a non-finite right-hand side
      rnorm = s->gam[0] * norm_fac;
      break;
    }
    int j = 0;  // columns completed in this cycle
    for (; j < m;) {
      ++iter;
      double* vj = s->V + (int64_t)j * n;
      double* zj = s->Z + (int64_t)j * n;
      double* vn = vj + n;
      // Z_j = M^-1 V_j and w = A Z_j, both sv_j times the normalised ones
      if ((rc = precond(*pc_kind, pc, n, zj, vj)) != kOk) return rc;
      if ((rc = apply_op(op, s->w, zj)) != kOk) return rc;
      // d_i = <w, V_i>_mult = sv_i sv_j h_ij, i <= j, one pass over w
      if (s->cnt_ok)
        mdots_launch(n, s->w, s->V, j + 1, (const uint8_t*)s->cnt8, s->part, s->G);
      else
        mdots_launch(n, s->w, s->V, j + 1, mult, s->part, s->G);
      hipLaunchKernelGGL(gm_fin_k, dim3(j + 1), dim3(kFinBlock), 0, stream(),
                         s->part, s->G, s->sc);
      // V_{j+1} = w - sum (d_i / nv_i) V_i = sv_j alpha v_{j+1}; its norm
      if (s->cnt_ok)
        orth(s->G, n, j + 1, s->sc, s->nv, s->V, s->w, vn, (const uint8_t*)s->cnt8,
             s->part);
      else
        orth(s->G, n, j + 1, s->sc, s->nv, s->V, s->w, vn, mult, s->part);
      hipLaunchKernelGGL(gm_report_k, dim3(1), dim3(kFinBlock), 0, stream(),
                         s->part, s->G, s->nv + j + 1, (const double*)s->sc,
                         j + 1, s->host);
      NK_LAUNCHED();
      NK_HIP(hipStreamSynchronize(stream()));
      const double svj = s->sv[j];
      for (int i = 0; i <= j; ++i) H(i, j) = s->host[i] / (s->sv[i] * svj);
      // previous rotations on column j
      for (int i = 0; i < j; ++i) {
        const double t = H(i, j);
        H(i, j) = s->c[i] * t + s->s[i] * H(i + 1, j);
        H(i + 1, j) = -s->s[i] * t + s->c[i] * H(i + 1, j);
      }
      const double svn = std::sqrt(std::max(s->host[j + 1], 0.0));
      const double alpha = svn / svj;
      if (alpha == 0.0) {  // lucky breakdown: the space is invariant
        conv = true;
        ++j;
        break;
      }
      const double lr = std::sqrt(H(j, j) * H(j, j) + alpha * alpha);
      s->c[j] = H(j, j) / lr;
      s->s[j] = alpha / lr;
      H(j, j) = lr;
      s->gam[j + 1] = -s->s[j] * s->gam[j];
      s->gam[j] = s->c[j] * s->gam[j];
      rnorm = std::fabs(s->gam[j + 1]) * norm_fac;
      ++j;
      if (rnorm < *abstol) {
        conv = true;
        break;
      }
      // a non-finite residual (a diverged flow upstream) ends the solve,
      // reported as not converged, instead of running to max_iter
      if (iter + 1 > *max_iter || !std::isfinite(rnorm)) break;
      s->sv[j] = svn;
      // the stored norms multiply by alpha each column: an unpreconditioned
      // operator (alpha ~ ||A||) would overflow nv within a long cycle, so a
      // vector far from unit length is scaled back to it (a rare pass)
      if (j < m && (svn > 1e60 || svn < 1e-60)) {
        hipLaunchKernelGGL(gm_scale_k, dim3(g), dim3(kGmBlock), 0, stream(),
                           n, vn, vn, 1.0 / svn);
        NK_LAUNCHED();
        if ((rc = mdot(s, vn, mult, s->nv + j)) != kOk) return rc;
        s->sv[j] = std::sqrt(s->host[0]);
      }
    }
    // back substitution H y = gam on the j completed columns, x += Z y / sv
    for (int k = j - 1; k >= 0; --k) {
      double t = s->gam[k];
      for (int i = j - 1; i > k; --i) t -= H(k, i) * s->y[i];
      s->y[k] = t / H(k, k);
    }
    // (the coefficients go with the launch, by value)
    GmCoef yc{};
    for (int k = 0; k < j; ++k) yc.c[k] = s->y[k] / s->sv[k];
    update(n, j, yc, s->Z, x, !xset);
    NK_LAUNCHED();
    xset = true;
    if (!std::isfinite(rnorm)) break;
  }
  if (!xset) {
    hipLaunchKernelGGL(gm_scale_k, dim3(g), dim3(kGmBlock), 0, stream(), n, x,
                       x, 0.0);
    NK_LAUNCHED();
  }
  s->mon.iter = iter;
  s->mon.done = conv ? 1 : 0;
  s->mon.res_final = rnorm;
  if (mon) *mon = s->mon;
  return kOk;
}

}  // extern "C"
