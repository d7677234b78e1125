// projection.hip -- solution projection for successive right-hand sides
// (Fischer, CMAME 163, 1998), the method of Neko's projection_t.
//
// Replaces proj_prs / proj_u / proj_v / proj_w of the reference
// (sources/adjoint/adjoint_pnpn.f90:102-105; init :365-373; pre_solving
// :729-730, :777-779; post_solving :739-740, :788-793).
//
// The basis lives in HBM as two L x n FP64 blocks, xx (A-orthogonal
// previous solutions u_i) and bb = A xx, with the squared A-norms
// nu_i^2 = <u_i, A u_i> kept as 1/nu_i^2 in a device array: the vectors are
// not rescaled to A-norm 1, every coefficient is divided by nu_i^2 instead
// (the same projection, one streaming pass fewer).  Per solve:
//   pre  : alpha_i = <b, u_i>_mult / nu_i^2, all i in one pass over b and
//          mult; then one pass xbar = sum alpha_i u_i, b -= sum alpha_i A u_i
//   post : x += xbar (+ the learned direction) and the new vector u_k in one
//          pass; A u_k (ax_helm + gs + bc); one pass for the Gram-Schmidt
//          coefficients and nu_k^2 before; one classical Gram-Schmidt pass in
//          the A-inner product that also forms nu_k^2 after; a second pass
//          only if the first removed more than half of the A-norm; the vector
//          joins the basis unless its A-norm fell below 1e-7 of the full
//          solution's.
// The new vector is the solver's correction dx = x - xbar (plus a learned
// direction), which spans the same space with the basis as x does: it is
// A-orthogonal to the basis up to the solve's residual, so one pass suffices
// where Neko's full x (|x| >> |dx| once the basis is good) cancels to |dx|
// and needs two.  After a restart xbar is no longer in the basis, and the
// full x is used.  ||x||_A for the dependence test follows from the
// coefficients: sum a_i^2 + 2 sum a_i d_i + ||dx||_A^2 (a: pre, d: first
// pass, both in the A-normalised basis).
// Coefficients stay on the device; the host reads them once per post (one
// copy of a few hundred bytes) to decide on the new vector and to record it.
//
// FP32 basis (nekoamd_proj_set_fp32): the previous solutions u_i are stored
// in FP32 and their images A u_i are not stored; all arithmetic stays FP64.
// pre forms xbar = sum alpha_i u_i and then b -= A xbar with one operator
// application (the operator recorded by the last post, its arrays as they
// are at the time of the call), so the projected right-hand side is exactly
// b - A xbar for the xbar handed back: the FP32 rounding of the basis only
// makes xbar a slightly different guess.  (Storing A u_i in FP32 instead
// would not do: the rounding of u_i is a rough field whose image under A is
// far larger than its own size.)  post forms A dx once, the Gram-Schmidt
// coefficients <A dx, u_i> / nu_i^2 from the FP32 basis, removes them from
// dx in FP64, stores the result in FP32 and takes its squared A-norm as
// ||dx||_A^2 - sum c_i^2 nu_i^2 (a second pass, with A applied again, when
// that removed more than three quarters of it).  Per point and solve that
// moves (m/2 + 2) + (m/2 + 1) + one operator for pre and (k/2 + 2) + (k/2 +
// 3) + one operator for post, in 8-byte units, against (3m + 5) and (3k +
// 13) + one operator with the FP64 basis.
//
// Restart that keeps the newest solutions (keep > 0; Neko's restart is
// keep = 0, an empty basis): the host records each accepted solution x_j by
// its coefficients in the A-normalised basis (column j of an upper
// triangular R: the two Gram-Schmidt passes' coefficients times nu_i, and
// nu_j on the diagonal).  When the basis is full, the last `keep` columns
// are factored R_q = Q T on the host (classical Gram-Schmidt, twice), and
// one in-place streaming pass per block forms u'_j = sum_i Q_ij u_i / nu_i
// (and A u'_j alike): A-orthonormal, spanning the kept solutions; R
// restarts as T.
// Algorithmic bytes per point (8 B units, basis size m, k = m before post):
//   pre (m + 2) + (2m + 3), post 5 + (k + 3) + (2k + 5) plus one operator
//   application (a second pass adds (k + 2) + (2k + 5)); dots hold one
//   accumulator per basis vector in registers.
#include <cmath>
#include <vector>

#include "internal.h"
#include "nekoamd.h"
#include "mdot.h"
#include "reduce.h"

namespace nekoamd {

constexpr double kProjTol = 1e-7;
constexpr int kDotGrid = 512;     // partials per inner product (augment)
constexpr int kProjGrid = kMultiGrid;  // max partials per inner product (basis)

struct ProjState {
  int64_t n = 0;
  int L = 0;
  int m = 0;
  int keep = 0;
  bool xbar_valid = false;
  std::vector<double> R;   // L x L, column-major: solution j's coefficients
  std::vector<double> nu;  // L: A-norm of the stored u_i
  double* xx = nullptr;    // L * n
  double* bb = nullptr;    // L * n
  double* xbar = nullptr;  // n: sum alpha_i u_i from pre
  double* xaug = nullptr;  // n: the learned directions added by augment
  bool xaug_valid = false;
  double* inv2 = nullptr;  // NEKOAMD_PROJ_MAX_L: 1 / nu_i^2
  // coef: [0, kCoefB) first-pass coefficients then nu^2 before,
  // [kCoefB, kCoefC) second-pass coefficients, kCoefC: nu^2 after,
  // kCoefC + 1..3: augment scratch, [kCoefA, kCoefN): the pre coefficients
  double* coef = nullptr;
  double* part = nullptr;  // (NEKOAMD_PROJ_MAX_L + 2) * kProjGrid
  double* cmat = nullptr;  // L x L: rebase coefficients
  double* host = nullptr;  // pinned: kCoefN read back, then L x L upload
  // FP32 basis: xx32 (L * n), the FP64 work vectors xw (the new vector) and
  // aw (operator images), and the operator of the last post
  bool fp32 = false;
  float* xx32 = nullptr;
  double* xw = nullptr;
  double* aw = nullptr;
  // the multiplicity weights as one-byte counts (mult = 1 / count exactly,
  // mult_counts), built on the first pre or post with a given mult array
  // (keyed by its address); the FP32-basis dot passes read them when cnt_ok
  uint8_t* cnt8 = nullptr;
  int* bad = nullptr;
  const void* cnt_key = nullptr;
  bool cnt_ok = false;
  bool op_valid = false;
  void* op_gs = nullptr;
  const int32_t* op_mask = nullptr;
  int64_t op_nmask = 0;
  const double *op_dx = nullptr, *op_h1 = nullptr, *op_h2 = nullptr,
               *op_g[6] = {}, *op_B = nullptr;
  int op_lx = 0, op_nelv = 0;
  bool op_ifh2 = false;
};

// Nontemporal load of the float pair p[0], p[1] (p 8-byte aligned).
__device__ __forceinline__ float2 ld_nt2f(const float* p) {
  const float2* q = (const float2*)p;
  float2 v;
  v.x = __builtin_nontemporal_load(&q->x);
  v.y = __builtin_nontemporal_load(&q->y);
  return v;
}

// part[i * G + blk] = block blk's partial of <a, V_i>_mult, i < m, for the m
// FP32 vectors V_i = V + i n (mdots_k with an FP32 basis).
template <int KM, bool P2, typename TM>
__global__ void __launch_bounds__(256)
    mdots32_k(int64_t n, const double* __restrict__ a,
              const float* __restrict__ V, int m,
              const TM* __restrict__ mult, double* __restrict__ part) {
  __shared__ double sh[KM][4];
  double s[KM];
#pragma unroll
  for (int i = 0; i < KM; ++i) s[i] = 0.0;
  if constexpr (P2) {
    const int64_t n2 = n / 2;
    const double2* a2 = (const double2*)a;
    for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < n2;
         q += (int64_t)gridDim.x * 256) {
      const double2 av = a2[q], mv = mult_pair(mult, q);
      const double w0 = mv.x * av.x, w1 = mv.y * av.y;
#pragma unroll
      for (int c = 0; c < KM; c += 8) {
        if (c < m) {
#pragma unroll
          for (int i = c; i < c + 8; ++i)
            if (i < m) {
              const float2 v = ld_nt2f(V + (int64_t)i * n + 2 * q);
              s[i] += w0 * (double)v.x + w1 * (double)v.y;
            }
          asm volatile("" ::: "memory");
        }
      }
    }
  } else {
    for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < n;
         p += (int64_t)gridDim.x * 256) {
      const double w = mult_at(mult, p) * a[p];
#pragma unroll
      for (int c = 0; c < KM; c += 8) {
        if (c < m) {
#pragma unroll
          for (int i = c; i < c + 8; ++i)
            if (i < m)
              s[i] += w * (double)__builtin_nontemporal_load(V + (int64_t)i * n + p);
          asm volatile("" ::: "memory");
        }
      }
    }
  }
  block_partials<KM>(s, m, sh, part);
}

// s += sum_{i<k} c[i] V_i (FP32 vectors) at the pair (2q, 2q + 1) (P2) or
// the point q (in s.x); c in LDS.
template <int KM, bool P2>
__device__ __forceinline__ void comb_point32(const double* c,
                                             const float* __restrict__ V,
                                             int64_t n, int k, int64_t q,
                                             double2& s) {
#pragma unroll
  for (int c0 = 0; c0 < KM; c0 += 8) {
    if (c0 < k) {
#pragma unroll
      for (int i = c0; i < c0 + 8; ++i)
        if (i < k) {
          const double ci = c[i];
          if constexpr (P2) {
            const float2 v = ld_nt2f(V + (int64_t)i * n + 2 * q);
            s.x += ci * (double)v.x;
            s.y += ci * (double)v.y;
          } else {
            s.x += ci * (double)__builtin_nontemporal_load(V + (int64_t)i * n + q);
          }
        }
      asm volatile("" ::: "memory");
    }
  }
}

// xout = sum_{i<m} alpha_i u_i over the FP32 basis.
template <int KM, bool P2>
__global__ void __launch_bounds__(256)
    proj_lincomb32_k(int64_t n, int m, const double* __restrict__ alpha,
                     const float* __restrict__ xx, double* __restrict__ xout) {
  __shared__ double sa[NEKOAMD_PROJ_MAX_L];
  if (threadIdx.x < m) sa[threadIdx.x] = alpha[threadIdx.x];
  __syncthreads();
  const int64_t nq = P2 ? n / 2 : n;
  for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < nq;
       q += (int64_t)gridDim.x * 256) {
    double2 sx = make_double2(0.0, 0.0);
    comb_point32<KM, P2>(sa, xx, n, m, q, sx);
    if constexpr (P2)
      ((double2*)xout)[q] = sx;
    else
      xout[q] = sx.x;
  }
}

// One Gram-Schmidt pass of xk against the first k FP32 basis vectors with
// coefficients c: xk -= sum c_i u_i (FP64), and its FP32 copy to uk.
template <int KM, bool P2>
__global__ void __launch_bounds__(256)
    proj_cgs32_k(int64_t n, int k, const double* __restrict__ c,
                 const float* __restrict__ xx, double* __restrict__ xk,
                 float* __restrict__ uk) {
  __shared__ double sc[KM];
  if ((int)threadIdx.x < k) sc[threadIdx.x] = c[threadIdx.x];
  __syncthreads();
  const int64_t nq = P2 ? n / 2 : n;
  for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < nq;
       q += (int64_t)gridDim.x * 256) {
    double2 sx = make_double2(0.0, 0.0);
    comb_point32<KM, P2>(sc, xx, n, k, q, sx);
    if constexpr (P2) {
      double2* x2 = (double2*)xk + q;
      const double2 xo = *x2;
      const double2 xn = make_double2(xo.x - sx.x, xo.y - sx.y);
      *x2 = xn;
      ((float2*)uk)[q] = make_float2((float)xn.x, (float)xn.y);
    } else {
      const double xn = xk[q] - sx.x;
      xk[q] = xn;
      uk[q] = (float)xn;
    }
  }
}

// In place, per point: V_j = sum_{i<m} C[i q + j] V_i for j < q (< m), FP32
// vectors, FP64 sums; every value of the point is read before the first write.
template <int KM>
__global__ void __launch_bounds__(256)
    proj_rebase32_k(int64_t n, int m, int q, const double* __restrict__ C,
                    float* V) {
  __shared__ double sC[KM * KM];
  for (int t = threadIdx.x; t < m * q; t += 256) sC[t] = C[t];
  __syncthreads();
  for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * 256) {
    double v[KM];
#pragma unroll
    for (int i = 0; i < KM; ++i)
      if (i < m) v[i] = (double)V[(int64_t)i * n + p];
#pragma unroll 1
    for (int j = 0; j < q; ++j) {
      double a = 0.0;
#pragma unroll
      for (int i = 0; i < KM; ++i)
        if (i < m) a += sC[i * q + j] * v[i];
      V[(int64_t)j * n + p] = (float)a;
    }
  }
}

// b -= a.
__global__ void __launch_bounds__(256)
    proj_sub_k(int64_t n, double* __restrict__ b, const double* __restrict__ a) {
  for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * 256)
    b[p] -= a[p];
}

// part[i * G + blk] = partial of sum_p mult a V_i over block blk's points.
template <typename TM>
__global__ void __launch_bounds__(256)
    proj_mdot_k(int64_t n, const double* __restrict__ a,
                const double* __restrict__ V, int64_t ldv,
                const TM* __restrict__ mult, double* __restrict__ part) {
  __shared__ double sh[256 / 64];
  const int i = blockIdx.y;
  const double* v = V + (int64_t)i * ldv;
  double s = 0.0;
  for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * 256)
    s += mult_at(mult, p) * a[p] * v[p];
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
template <int KM, bool P2>
__global__ void __launch_bounds__(256)
    proj_lincomb_k(int64_t n, int m, const double* __restrict__ alpha,
                   const double* __restrict__ xx,
                   const double* __restrict__ bb, double* __restrict__ xout,
                   double* __restrict__ b, int assign) {
  __shared__ double sa[NEKOAMD_PROJ_MAX_L];
  if (threadIdx.x < m) sa[threadIdx.x] = alpha[threadIdx.x];
  __syncthreads();
  const int64_t nq = P2 ? n / 2 : n;
  for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < nq;
       q += (int64_t)gridDim.x * 256) {
    double2 sx = make_double2(0.0, 0.0), sb = make_double2(0.0, 0.0);
    comb_point<KM, P2, true>(sa, xx, bb, n, m, q, sx, sb);
    if constexpr (P2) {
      double2* x2 = (double2*)xout + q;
      double2* b2 = (double2*)b + q;
      const double2 bo = *b2;
      if (assign) {
        *x2 = sx;
      } else {
        const double2 xo = *x2;
        *x2 = make_double2(xo.x - sx.x, xo.y - sx.y);
      }
      *b2 = make_double2(bo.x - sb.x, bo.y - sb.y);
    } else {
      xout[q] = assign ? sx.x : xout[q] - sx.x;
      b[q] -= sb.x;
    }
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

constexpr int kCoefB = NEKOAMD_PROJ_MAX_L + 1;
constexpr int kCoefC = 2 * NEKOAMD_PROJ_MAX_L + 2;
constexpr int kCoefA = kCoefC + 4;
constexpr int kCoefN = kCoefA + NEKOAMD_PROJ_MAX_L;
constexpr double kReorth = 0.5;  // second pass if ||v'||_A < 0.5 ||v||_A

// One classical Gram-Schmidt pass of (u_k, b_k) = (xk, bk) against the first
// k basis vectors with coefficients c: xk -= sum c_i u_i, bk -= sum c_i A u_i,
// with the partials of <xk', bk'>_mult (the squared A-norm after the pass)
// from the same sweep.
template <int KM, bool P2>
__global__ void __launch_bounds__(256)
    proj_cgs_k(int64_t n, int k, const double* __restrict__ c,
               const double* __restrict__ xx, const double* __restrict__ bb,
               double* __restrict__ xk, double* __restrict__ bk,
               const double* __restrict__ mult, double* __restrict__ part) {
  __shared__ double sc[KM];
  __shared__ double sh[1][4];
  if ((int)threadIdx.x < k) sc[threadIdx.x] = c[threadIdx.x];
  __syncthreads();
  double s[1] = {0.0};
  const int64_t nq = P2 ? n / 2 : n;
  for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < nq;
       q += (int64_t)gridDim.x * 256) {
    double2 sx = make_double2(0.0, 0.0), sb = make_double2(0.0, 0.0);
    comb_point<KM, P2, true>(sc, xx, bb, n, k, q, sx, sb);
    if constexpr (P2) {
      double2* x2 = (double2*)xk + q;
      double2* b2 = (double2*)bk + q;
      const double2 xo = *x2, bo = *b2, mv = ((const double2*)mult)[q];
      const double2 xn = make_double2(xo.x - sx.x, xo.y - sx.y);
      const double2 bn = make_double2(bo.x - sb.x, bo.y - sb.y);
      *x2 = xn;
      *b2 = bn;
      s[0] += mv.x * xn.x * bn.x + mv.y * xn.y * bn.y;
    } else {
      const double xn = xk[q] - sx.x, bn = bk[q] - sb.x;
      xk[q] = xn;
      bk[q] = bn;
      s[0] += mult[q] * xn * bn;
    }
  }
  block_partials<1>(s, 1, sh, part);
}

// out[i] = (sum of the G partials of i) / nu_i^2 for i < ns, unscaled from
// ns on; one block per i.
__global__ void __launch_bounds__(kFinBlock)
    proj_fin_k(const double* __restrict__ part, int G,
               const double* __restrict__ inv2, int ns,
               double* __restrict__ out) {
  __shared__ double sh[kFinBlock / 64];
  const int i = blockIdx.x;
  const double s = sum_partials<kFinBlock>(part + (int64_t)i * G, G, sh);
  if (threadIdx.x == 0) out[i] = i < ns ? s * inv2[i] : s;
}

// x_out = x + [xbar] + [xaug]; u_k = x_out (full) or x + [xaug] (the
// correction and the learned direction).
__global__ void __launch_bounds__(256)
    proj_restore_k(int64_t n, double* __restrict__ x,
                   const double* __restrict__ xbar,
                   const double* __restrict__ xaug, double* __restrict__ xk,
                   int full) {
  for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * 256) {
    const double xi = x[p];
    const double a = xaug ? xaug[p] : 0.0;
    double xo = xi;
    if (xbar) xo += xbar[p];
    if (xaug) xo += a;
    if (xbar || xaug) x[p] = xo;
    xk[p] = full ? xo : (xaug ? xi + a : xi);
  }
}

__global__ void proj_setinv_k(double* __restrict__ inv2, int k,
                              const double* __restrict__ nu2) {
  if (threadIdx.x == 0) inv2[k] = 1.0 / nu2[0];
}

__global__ void proj_fill_k(double* __restrict__ a, int q, double v) {
  if ((int)threadIdx.x < q) a[threadIdx.x] = v;
}

// In place, per point: V_j = sum_{i<m} C[i q + j] V_i for j < q (< m); every
// value of the point is read before the first write.
template <int KM>
__global__ void __launch_bounds__(256)
    proj_rebase_k(int64_t n, int m, int q, const double* __restrict__ C,
                  double* V) {
  __shared__ double sC[KM * KM];
  for (int t = threadIdx.x; t < m * q; t += 256) sC[t] = C[t];
  __syncthreads();
  for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * 256) {
    double v[KM];
#pragma unroll
    for (int i = 0; i < KM; ++i)
      if (i < m) v[i] = V[(int64_t)i * n + p];
#pragma unroll 1
    for (int j = 0; j < q; ++j) {
      double a = 0.0;
#pragma unroll
      for (int i = 0; i < KM; ++i)
        if (i < m) a += sC[i * q + j] * v[i];
      V[(int64_t)j * n + p] = a;
    }
  }
}

static ProjState* as_proj(void* p) { return (ProjState*)p; }

// The one-byte counts of mult for the FP32 basis's dot passes, built once
// per mult array (synchronises when it builds).
static int ensure_counts(ProjState* s, const double* mult) {
  if (!s->fp32 || s->n == 0 || mult == s->cnt_key) return kOk;
  if (!s->cnt8) NK_HIP(hipMalloc((void**)&s->cnt8, (size_t)s->n));
  if (!s->bad) NK_HIP(hipMalloc((void**)&s->bad, sizeof(int)));
  const int rc = mult_counts(s->n, mult, s->cnt8, s->bad, (int*)s->host,
                             &s->cnt_ok);
  if (rc) return rc;
  s->cnt_key = mult;
  return kOk;
}

// out[0..m) = <a, V_i>_mult for the m vectors V_i = V + i * ldv.
static int mdot(ProjState* s, const double* a, const double* V, int64_t ldv,
                int m, const double* mult, double* out) {
  if (m <= 0) return kOk;
  const int G = nk_grid(s->n, 256, kDotGrid);
  if (s->cnt_ok && mult == s->cnt_key)
    hipLaunchKernelGGL(proj_mdot_k<uint8_t>, dim3(G, m), dim3(256), 0, stream(),
                       s->n, a, V, ldv, (const uint8_t*)s->cnt8, s->part);
  else
    hipLaunchKernelGGL(proj_mdot_k<double>, dim3(G, m), dim3(256), 0, stream(),
                       s->n, a, V, ldv, mult, s->part);
  NK_LAUNCHED();
  hipLaunchKernelGGL(proj_mdot_fin_k, dim3(m), dim3(kFinBlock), 0, stream(),
                     s->part, G, out);
  NK_LAUNCHED();
  return kOk;
}

// out[0..m) = <a, u_i>_mult for the m basis vectors from V (stride n),
// entries i < ns divided by nu_i^2.
template <int KM>
static void cgs_launch(ProjState* s, int k, const double* c, double* xk,
                       double* bk, const double* mult, int G, bool p2) {
  const dim3 g(G), b(256);
  if (p2)
    hipLaunchKernelGGL((proj_cgs_k<KM, true>), g, b, 0, stream(), s->n, k, c,
                       s->xx, s->bb, xk, bk, mult, s->part);
  else
    hipLaunchKernelGGL((proj_cgs_k<KM, false>), g, b, 0, stream(), s->n, k, c,
                       s->xx, s->bb, xk, bk, mult, s->part);
}

template <int KM>
static void lincomb_launch(ProjState* s, int m, const double* alpha,
                           double* xout, double* b, int assign, bool p2) {
  const dim3 g(kMultiGrid), bl(256);
  if (p2)
    hipLaunchKernelGGL((proj_lincomb_k<KM, true>), g, bl, 0, stream(), s->n, m,
                       alpha, s->xx, s->bb, xout, b, assign);
  else
    hipLaunchKernelGGL((proj_lincomb_k<KM, false>), g, bl, 0, stream(), s->n,
                       m, alpha, s->xx, s->bb, xout, b, assign);
}

// xout = [xout -] sum alpha_i u_i, b -= sum alpha_i A u_i over the first m
// basis vectors, one pass.
static void lincomb(ProjState* s, int m, const double* alpha, double* xout,
                    double* b, int assign) {
  const bool p2 = pairs_ok(s->n, s->xx, s->bb, xout, b);
  if (m <= 8)
    lincomb_launch<8>(s, m, alpha, xout, b, assign, p2);
  else if (m <= 16)
    lincomb_launch<16>(s, m, alpha, xout, b, assign, p2);
  else if (m <= 32)
    lincomb_launch<32>(s, m, alpha, xout, b, assign, p2);
  else
    lincomb_launch<64>(s, m, alpha, xout, b, assign, p2);
}

static int dots(ProjState* s, const double* a, const double* V, int m, int ns,
                const double* mult, double* out) {
  if (m <= 0) return kOk;
  const int G = nk_grid(s->n, 256, kProjGrid);
  mdots_launch(s->n, a, V, m, mult, s->part, G);
  NK_LAUNCHED();
  hipLaunchKernelGGL(proj_fin_k, dim3(m), dim3(kFinBlock), 0, stream(),
                     s->part, G, s->inv2, ns, out);
  NK_LAUNCHED();
  return kOk;
}

// One Gram-Schmidt pass against the first k vectors with coefficients c;
// out[0] receives nu_k^2 after.
static int cgs(ProjState* s, int k, const double* c, double* xk, double* bk,
               const double* mult, double* out) {
  const int G = nk_grid(s->n, 256, kProjGrid);
  const bool p2 = pairs_ok(s->n, s->xx, s->bb, xk, bk, mult);
  if (k <= 8)
    cgs_launch<8>(s, k, c, xk, bk, mult, G, p2);
  else if (k <= 16)
    cgs_launch<16>(s, k, c, xk, bk, mult, G, p2);
  else if (k <= 32)
    cgs_launch<32>(s, k, c, xk, bk, mult, G, p2);
  else
    cgs_launch<64>(s, k, c, xk, bk, mult, G, p2);
  NK_LAUNCHED();
  hipLaunchKernelGGL(proj_fin_k, dim3(1), dim3(kFinBlock), 0, stream(),
                     s->part, G, s->inv2, 0, out);
  NK_LAUNCHED();
  return kOk;
}

// ---- FP32 basis --------------------------------------------------------
template <int KM, typename TM>
static void mdots32_launch_km(ProjState* s, const double* a, int m,
                              const TM* mult, int G, bool p2) {
  const dim3 g(G), b(256);
  if (p2)
    hipLaunchKernelGGL((mdots32_k<KM, true, TM>), g, b, 0, stream(), s->n, a,
                       s->xx32, m, mult, s->part);
  else
    hipLaunchKernelGGL((mdots32_k<KM, false, TM>), g, b, 0, stream(), s->n, a,
                       s->xx32, m, mult, s->part);
}

template <typename TM>
static void mdots32_launch(ProjState* s, const double* a, int m, const TM* mult,
                           int G, bool p2) {
  if (m <= 8)
    mdots32_launch_km<8>(s, a, m, mult, G, p2);
  else if (m <= 16)
    mdots32_launch_km<16>(s, a, m, mult, G, p2);
  else if (m <= 24)
    mdots32_launch_km<24>(s, a, m, mult, G, p2);
  else if (m <= 32)
    mdots32_launch_km<32>(s, a, m, mult, G, p2);
  else
    mdots32_launch_km<64>(s, a, m, mult, G, p2);
}

// out[0..m) = <a, u_i>_mult / nu_i^2 for the first m FP32 basis vectors.
static int dots32(ProjState* s, const double* a, int m, const double* mult,
                  double* out) {
  if (m <= 0) return kOk;
  const int G = nk_grid(s->n, 256, kProjGrid);
  if (s->cnt_ok && mult == s->cnt_key) {
    const bool p2 = pairs_ok(s->n, a, s->xx32) && ((uintptr_t)s->cnt8 & 1) == 0;
    mdots32_launch(s, a, m, (const uint8_t*)s->cnt8, G, p2);
  } else {
    mdots32_launch(s, a, m, mult, G, pairs_ok(s->n, a, mult, s->xx32));
  }
  NK_LAUNCHED();
  hipLaunchKernelGGL(proj_fin_k, dim3(m), dim3(kFinBlock), 0, stream(),
                     s->part, G, s->inv2, m, out);
  NK_LAUNCHED();
  return kOk;
}

template <int KM>
static void lincomb32_launch(ProjState* s, int m, const double* alpha,
                             double* xout, bool p2) {
  const dim3 g(kMultiGrid), bl(256);
  if (p2)
    hipLaunchKernelGGL((proj_lincomb32_k<KM, true>), g, bl, 0, stream(), s->n,
                       m, alpha, s->xx32, xout);
  else
    hipLaunchKernelGGL((proj_lincomb32_k<KM, false>), g, bl, 0, stream(),
                       s->n, m, alpha, s->xx32, xout);
}

// xout = sum_{i<m} alpha_i u_i over the FP32 basis.
static int lincomb32(ProjState* s, int m, const double* alpha, double* xout) {
  const bool p2 = pairs_ok(s->n, xout, s->xx32);
  if (m <= 8)
    lincomb32_launch<8>(s, m, alpha, xout, p2);
  else if (m <= 16)
    lincomb32_launch<16>(s, m, alpha, xout, p2);
  else if (m <= 32)
    lincomb32_launch<32>(s, m, alpha, xout, p2);
  else
    lincomb32_launch<64>(s, m, alpha, xout, p2);
  NK_LAUNCHED();
  return kOk;
}

template <int KM>
static void cgs32_launch(ProjState* s, int k, const double* c, double* xk,
                         float* uk, bool p2) {
  const dim3 g(kMultiGrid), bl(256);
  if (p2)
    hipLaunchKernelGGL((proj_cgs32_k<KM, true>), g, bl, 0, stream(), s->n, k,
                       c, s->xx32, xk, uk);
  else
    hipLaunchKernelGGL((proj_cgs32_k<KM, false>), g, bl, 0, stream(), s->n, k,
                       c, s->xx32, xk, uk);
}

// xk -= sum_{i<k} c_i u_i (k may be 0) and uk = (float) xk.
static int cgs32(ProjState* s, int k, const double* c, double* xk, float* uk) {
  const bool p2 = pairs_ok(s->n, xk, uk, s->xx32);
  if (k <= 8)
    cgs32_launch<8>(s, k, c, xk, uk, p2);
  else if (k <= 16)
    cgs32_launch<16>(s, k, c, xk, uk, p2);
  else if (k <= 32)
    cgs32_launch<32>(s, k, c, xk, uk, p2);
  else
    cgs32_launch<64>(s, k, c, xk, uk, p2);
  NK_LAUNCHED();
  return kOk;
}

// aw = bc(gs(A v)) with the operator recorded by the last post.
static int apply_op(ProjState* s, const double* v) {
  int rc = ax_helm_launch<double, double>(
      s->aw, v, s->op_dx, s->op_h1, s->op_h2, s->op_g[0], s->op_g[1],
      s->op_g[2], s->op_g[3], s->op_g[4], s->op_g[5], s->op_B, s->op_lx,
      s->op_nelv, s->op_ifh2, nullptr);
  if (!rc) rc = gs_launch<double>(s->op_gs, s->aw, NEKOAMD_GS_ADD, nullptr);
  if (!rc) rc = bc_zero_launch<double>(s->aw, s->op_mask, s->op_nmask, nullptr);
  return rc;
}

// Thin QR of the m x q column-major A (full column rank expected) by
// classical Gram-Schmidt with one reorthogonalisation: Q (m x q, orthonormal
// columns), T = Q^T A (q x q, column-major).  Returns the number of columns
// kept: a column that is dependent on the previous ones to 1e-14 is dropped.
static int thin_qr(int m, int q, const double* A, double* Q, double* T) {
  int r = 0;
  std::vector<double> v(m), c(q);
  for (int j = 0; j < q; ++j) {
    const double* a = A + (int64_t)j * m;
    double na = 0.0;
    for (int i = 0; i < m; ++i) {
      v[i] = a[i];
      na += a[i] * a[i];
    }
    for (int pass = 0; pass < 2; ++pass)
      for (int l = 0; l < r; ++l) {
        const double* ql = Q + (int64_t)l * m;
        double d = 0.0;
        for (int i = 0; i < m; ++i) d += ql[i] * v[i];
        for (int i = 0; i < m; ++i) v[i] -= d * ql[i];
      }
    double nv = 0.0;
    for (int i = 0; i < m; ++i) nv += v[i] * v[i];
    if (!(nv > 1e-28 * na)) continue;
    const double s = 1.0 / std::sqrt(nv);
    double* qr = Q + (int64_t)r * m;
    for (int i = 0; i < m; ++i) qr[i] = v[i] * s;
    ++r;
  }
  for (int j = 0; j < q; ++j)
    for (int l = 0; l < q; ++l) {
      double d = 0.0;
      if (l < r)
        for (int i = 0; i < m; ++i) d += Q[(int64_t)l * m + i] * A[(int64_t)j * m + i];
      T[l + (int64_t)q * j] = d;
    }
  return r;
}

// The full basis restarts from the span of the `keep` newest solutions.
static int rebase(ProjState* s) {
  const int m = s->m, L = s->L, q = s->keep;
  if (q == 0) {
    s->m = 0;
    return kOk;
  }
  std::vector<double> A((size_t)m * q), Q((size_t)m * q), T((size_t)q * q);
  for (int j = 0; j < q; ++j)
    for (int i = 0; i < m; ++i)
      A[(size_t)j * m + i] = s->R[(size_t)(m - q + j) * L + i];
  const int r = thin_qr(m, q, A.data(), Q.data(), T.data());
  if (r == 0) {
    s->m = 0;
    return kOk;
  }
  // C[i r + j] = Q_ij / nu_i (u_i are stored with A-norm nu_i)
  double* C = s->host + kCoefN;
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < r; ++j) C[i * r + j] = Q[(size_t)j * m + i] / s->nu[i];
  NK_HIP(hipMemcpyAsync(s->cmat, C, sizeof(double) * m * r,
                        hipMemcpyHostToDevice, stream()));
  const int G = nk_grid(s->n, 256);
  if (s->fp32) {
    if (L <= 8)
      hipLaunchKernelGGL(proj_rebase32_k<8>, dim3(G), dim3(256), 0, stream(),
                         s->n, m, r, s->cmat, s->xx32);
    else if (L <= 16)
      hipLaunchKernelGGL(proj_rebase32_k<16>, dim3(G), dim3(256), 0, stream(),
                         s->n, m, r, s->cmat, s->xx32);
    else if (L <= 32)
      hipLaunchKernelGGL(proj_rebase32_k<32>, dim3(G), dim3(256), 0, stream(),
                         s->n, m, r, s->cmat, s->xx32);
    else
      hipLaunchKernelGGL(proj_rebase32_k<64>, dim3(G), dim3(256), 0, stream(),
                         s->n, m, r, s->cmat, s->xx32);
    NK_LAUNCHED();
  }
  for (double* V : {s->xx, s->bb}) {
    if (s->fp32) break;
    if (L <= 8)
      hipLaunchKernelGGL(proj_rebase_k<8>, dim3(G), dim3(256), 0, stream(),
                         s->n, m, r, s->cmat, V);
    else if (L <= 16)
      hipLaunchKernelGGL(proj_rebase_k<16>, dim3(G), dim3(256), 0, stream(),
                         s->n, m, r, s->cmat, V);
    else if (L <= 32)
      hipLaunchKernelGGL(proj_rebase_k<32>, dim3(G), dim3(256), 0, stream(),
                         s->n, m, r, s->cmat, V);
    else
      hipLaunchKernelGGL(proj_rebase_k<64>, dim3(G), dim3(256), 0, stream(),
                         s->n, m, r, s->cmat, V);
    NK_LAUNCHED();
  }
  hipLaunchKernelGGL(proj_fill_k, dim3(1), dim3(64), 0, stream(), s->inv2, r,
                     1.0);
  NK_LAUNCHED();
  std::fill(s->R.begin(), s->R.end(), 0.0);
  for (int j = 0; j < r; ++j) {
    s->nu[j] = 1.0;
    for (int l = 0; l <= j; ++l) s->R[(size_t)j * L + l] = T[l + (size_t)q * j];
  }
  s->m = r;
  return kOk;
}

// post with the FP32 basis (see the file header): the new vector k = s->m
// is the correction (or the full solution when `full`) in s->xw.
static int post32(ProjState* s, double* X, const double* M, bool had_pre,
                  int mpre, bool full) {
  const int64_t n = s->n;
  const int ge = nk_grid(n, 256);
  const int k = s->m;
  double* xk = s->xw;
  float* uk = s->xx32 + (int64_t)k * n;
  hipLaunchKernelGGL(proj_restore_k, dim3(ge), dim3(256), 0, stream(), n, X,
                     had_pre ? s->xbar : nullptr,
                     s->xaug_valid ? s->xaug : nullptr, xk, full ? 1 : 0);
  NK_LAUNCHED();
  s->xbar_valid = false;
  s->xaug_valid = false;
  double* c1 = s->coef;
  double* c2 = s->coef + kCoefB;
  // c[i] = <A xk, u_i> / nu_i^2 (i < k), c[k] = ||xk||_A^2; xk -= sum c_i u_i
  auto pass = [&](double* c) {
    int rc = apply_op(s, xk);
    if (!rc) rc = dots32(s, s->aw, k, M, c);
    if (!rc) rc = mdot(s, xk, s->aw, 0, 1, M, c + k);
    if (!rc) rc = cgs32(s, k, c, xk, uk);
    return rc;
  };
  int rc = pass(c1);
  if (rc) return rc;
  if (const int rc_ = to_host(s->host, s->coef, sizeof(double) * kCoefN))
    return rc_;
  NK_HIP(hipStreamSynchronize(stream()));
  const double* h = s->host;
  const double v2 = h[k];
  // ||xk - sum c_i u_i||_A^2 = ||xk||_A^2 - sum (c_i nu_i)^2 (A-orthogonal u_i)
  auto after = [&](const double* c) {
    double r = c[k];
    for (int i = 0; i < k; ++i) r -= (c[i] * s->nu[i]) * (c[i] * s->nu[i]);
    return r;
  };
  double a1 = after(h);
  std::vector<double> col(s->L, 0.0);
  double full2 = v2;
  if (!full) {
    double aa = 0.0, ad = 0.0;
    for (int i = 0; i < k && i < mpre; ++i) {
      const double a = h[kCoefA + i] * s->nu[i], d = h[i] * s->nu[i];
      aa += a * a;
      ad += a * d;
      col[i] = a;
    }
    full2 = aa + 2.0 * ad + v2;
  }
  for (int i = 0; i < k; ++i) col[i] += h[i] * s->nu[i];
  if (k > 0 && a1 < kReorth * kReorth * v2) {
    // the first pass cancelled: a second one, with A applied again
    rc = pass(c2);
    if (rc) return rc;
    if (const int rc_ = to_host(s->host + kCoefB, c2, sizeof(double) * (k + 1)))
      return rc_;
    NK_HIP(hipStreamSynchronize(stream()));
    for (int i = 0; i < k; ++i) col[i] += h[kCoefB + i] * s->nu[i];
    a1 = after(h + kCoefB);
  }
  const double n1 = sqrt(a1 > 0.0 ? a1 : 0.0);
  const double nf = sqrt(full2 > 0.0 ? full2 : 0.0);
  if (n1 > kProjTol * nf && n1 > 0.0) {
    hipLaunchKernelGGL(proj_fill_k, dim3(1), dim3(64), 0, stream(),
                       s->inv2 + k, 1, 1.0 / a1);
    NK_LAUNCHED();
    double* R = s->R.data() + (size_t)k * s->L;
    std::copy(col.begin(), col.end(), R);
    R[k] = n1;
    s->nu[k] = n1;
    s->m = k + 1;
  }
  return kOk;
}

}  // namespace nekoamd

using namespace nekoamd;

extern "C" {

int nekoamd_proj_set_fp32(void* proj, const int* on) {
  NK_INIT();
  ProjState* s = as_proj(proj);
  NK_REQUIRE(s && on, "null argument");
  NK_REQUIRE(s->m == 0, "the basis must be empty");
  const bool f = *on != 0;
  if (f == s->fp32) return kOk;
  NK_HIP(hipStreamSynchronize(stream()));
  const int64_t N = s->n > 0 ? s->n : 1;
  hipError_t e = hipSuccess;
  if (f) {
    NK_HIP(hipFree(s->xx));
    NK_HIP(hipFree(s->bb));
    s->xx = s->bb = nullptr;
    e = hipMalloc((void**)&s->xx32, sizeof(float) * N * s->L);
    if (e == hipSuccess) e = hipMalloc((void**)&s->xw, sizeof(double) * N);
    if (e == hipSuccess) e = hipMalloc((void**)&s->aw, sizeof(double) * N);
  } else {
    NK_HIP(hipFree(s->xx32));
    NK_HIP(hipFree(s->xw));
    NK_HIP(hipFree(s->aw));
    s->xx32 = nullptr;
    s->xw = s->aw = nullptr;
    e = hipMalloc((void**)&s->xx, sizeof(double) * N * s->L);
    if (e == hipSuccess) e = hipMalloc((void**)&s->bb, sizeof(double) * N * s->L);
  }
  if (e != hipSuccess) return fail_hip(__func__, e);
  s->fp32 = f;
  return kOk;
}

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
  if (e == hipSuccess) e = hipMalloc((void**)&s->xaug, sizeof(double) * N);
  if (e == hipSuccess)
    e = hipMalloc((void**)&s->inv2, sizeof(double) * NEKOAMD_PROJ_MAX_L);
  if (e == hipSuccess)
    e = hipMalloc((void**)&s->coef, sizeof(double) * kCoefN);
  if (e == hipSuccess)
    e = hipMalloc((void**)&s->part,
                  sizeof(double) * (NEKOAMD_PROJ_MAX_L + 2) * kProjGrid);
  if (e == hipSuccess)
    e = hipMalloc((void**)&s->cmat, sizeof(double) * s->L * s->L);
  if (e == hipSuccess)
    e = hipHostMalloc((void**)&s->host,
                      sizeof(double) * (kCoefN + s->L * s->L),
                      hipHostMallocDefault);
  s->R.assign((size_t)s->L * s->L, 0.0);
  s->nu.assign(s->L, 1.0);
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
  void* dev[] = {s->xx,   s->bb,   s->xbar, s->xaug, s->inv2, s->coef,
                 s->part, s->cmat, s->xx32, s->xw,   s->aw,   s->cnt8,
                 s->bad};
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
  s->xaug_valid = false;
  return kOk;
}

int nekoamd_proj_size(void* proj, int* m) {
  ProjState* s = as_proj(proj);
  if (!s || !m) return fail_arg(__func__, "null argument");
  *m = s->m;
  return kOk;
}

int nekoamd_proj_set_keep(void* proj, const int* keep) {
  ProjState* s = as_proj(proj);
  if (!s || !keep) return fail_arg(__func__, "null argument");
  if (*keep < 0 || *keep >= s->L)
    return fail_arg(__func__, "keep must be in [0, L - 1]");
  s->keep = *keep;
  return kOk;
}

int nekoamd_proj_pre(void* proj, void* b, const void* mult) {
  NK_INIT();
  ProjState* s = as_proj(proj);
  NK_REQUIRE(s && (s->n == 0 || (b && mult)), "null argument");
  s->xbar_valid = false;
  s->xaug_valid = false;
  if (s->m == 0 || s->n == 0) return kOk;
  double* alpha = s->coef + kCoefA;
  if (s->fp32) {
    // xbar = sum alpha_i u_i; b -= A xbar with the last post's operator
    NK_REQUIRE(s->op_valid, "no operator recorded by a post");
    int rc = ensure_counts(s, (const double*)mult);
    if (!rc) rc = dots32(s, (const double*)b, s->m, (const double*)mult, alpha);
    if (!rc) rc = lincomb32(s, s->m, alpha, s->xbar);
    if (!rc) rc = apply_op(s, s->xbar);
    if (rc) return rc;
    hipLaunchKernelGGL(proj_sub_k, dim3(nk_grid(s->n, 256)), dim3(256), 0,
                       stream(), s->n, (double*)b, s->aw);
    NK_LAUNCHED();
    s->xbar_valid = true;
    return kOk;
  }
  int rc = dots(s, (const double*)b, s->xx, s->m, s->m, (const double*)mult,
                alpha);
  if (rc) return rc;
  lincomb(s, s->m, alpha, s->xbar, (double*)b, 1);
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
  double* qd = s->coef + kCoefC + 1;
  int rc = mdot(s, (const double*)q, (const double*)b, 0, 1,
                (const double*)mult, qd);
  if (!rc)
    rc = mdot(s, (const double*)q, (const double*)aq, 0, 1,
              (const double*)mult, qd + 1);
  if (rc) return rc;
  hipLaunchKernelGGL(proj_augment_k, dim3(nk_grid(s->n, 256)), dim3(256), 0,
                     stream(), s->n, (const double*)q, (const double*)aq,
                     s->xaug, (double*)b, qd, s->xaug_valid ? 0 : 1);
  NK_LAUNCHED();
  s->xaug_valid = true;
  if (const int rc_ = to_host(s->host, qd + 2, sizeof(double))) return rc_;
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
  // the operator, for pre with the FP32 basis
  s->op_valid = true;
  s->op_gs = gs;
  s->op_mask = (const int32_t*)mask;
  s->op_nmask = *nmask;
  s->op_dx = (const double*)dx;
  s->op_h1 = (const double*)h1;
  s->op_h2 = (const double*)h2;
  const void* gg[6] = {g11, g12, g13, g22, g23, g33};
  for (int q = 0; q < 6; ++q) s->op_g[q] = (const double*)gg[q];
  s->op_B = (const double*)B;
  s->op_lx = *lx;
  s->op_nelv = *nelv;
  s->op_ifh2 = *ifh2 != 0;
  const int64_t n = s->n;
  const int ge = nk_grid(n, 256);
  double* X = (double*)x;
  const double* M = (const double*)mult;
  if (const int rc0 = ensure_counts(s, M)) return rc0;
  const bool had_pre = s->xbar_valid;
  const int mpre = s->m;
  bool restarted = false;
  if (s->m == s->L) {
    const int rc = rebase(s);
    if (rc) return rc;
    restarted = true;
  }
  // the full solution after a restart (xbar left the basis) or without a
  // pre; else the correction (+ the learned direction)
  const bool full = restarted || !had_pre;
  if (s->fp32) return post32(s, X, M, had_pre, mpre, full);
  const int k = s->m;
  double* xk = s->xx + (int64_t)k * n;
  double* bk = s->bb + (int64_t)k * n;
  hipLaunchKernelGGL(proj_restore_k, dim3(ge), dim3(256), 0, stream(), n, X,
                     had_pre ? s->xbar : nullptr,
                     s->xaug_valid ? s->xaug : nullptr, xk, full ? 1 : 0);
  NK_LAUNCHED();
  s->xbar_valid = false;
  s->xaug_valid = false;
  int rc = ax_helm_launch<double, double>(
      bk, xk, (const double*)dx, (const double*)h1, (const double*)h2,
      (const double*)g11, (const double*)g12, (const double*)g13,
      (const double*)g22, (const double*)g23, (const double*)g33,
      (const double*)B, *lx, *nelv, *ifh2 != 0, nullptr);
  if (!rc) rc = gs_launch<double>(gs, bk, NEKOAMD_GS_ADD, nullptr);
  if (!rc)
    rc = bc_zero_launch<double>(bk, (const int32_t*)mask, *nmask, nullptr);
  if (rc) return rc;
  // coef[0, k): first-pass coefficients; coef[k] = nu_k^2 before
  double* c1 = s->coef;
  double* c2 = s->coef + kCoefB;
  double* nu2 = s->coef + kCoefC;
  rc = dots(s, xk, s->bb, k + 1, k, M, c1);
  if (!rc && k > 0) rc = cgs(s, k, c1, xk, bk, M, nu2);
  if (rc) return rc;
  if (const int rc_ = to_host(s->host, s->coef, sizeof(double) * kCoefN))
    return rc_;
  NK_HIP(hipStreamSynchronize(stream()));
  const double* h = s->host;
  const double v2 = h[k];
  double a1 = k > 0 ? h[kCoefC] : v2;
  // coefficients of x_out in the A-normalised basis, and ||x_out||_A
  std::vector<double> col(s->L, 0.0);
  double full2 = v2;
  if (!full) {
    double aa = 0.0, ad = 0.0;
    for (int i = 0; i < k && i < mpre; ++i) {
      const double a = h[kCoefA + i] * s->nu[i], d = h[i] * s->nu[i];
      aa += a * a;
      ad += a * d;
      col[i] = a;
    }
    full2 = aa + 2.0 * ad + v2;
  }
  for (int i = 0; i < k; ++i) col[i] += h[i] * s->nu[i];
  if (k > 0 && a1 < kReorth * kReorth * v2) {
    // the first pass cancelled: a second one
    rc = dots(s, xk, s->bb, k, k, M, c2);
    if (!rc) rc = cgs(s, k, c2, xk, bk, M, nu2);
    if (rc) return rc;
    if (const int rc_ = to_host(s->host + kCoefB, c2,
                                sizeof(double) * (kCoefC + 1 - kCoefB)))
      return rc_;
    NK_HIP(hipStreamSynchronize(stream()));
    for (int i = 0; i < k; ++i) col[i] += h[kCoefB + i] * s->nu[i];
    a1 = h[kCoefC];
  }
  const double n1 = sqrt(a1 > 0.0 ? a1 : 0.0);
  const double nf = sqrt(full2 > 0.0 ? full2 : 0.0);
  if (n1 > kProjTol * nf && n1 > 0.0) {
    hipLaunchKernelGGL(proj_setinv_k, dim3(1), dim3(64), 0, stream(), s->inv2,
                       k, k > 0 ? nu2 : c1 + k);
    NK_LAUNCHED();
    // record x_out = sum_{i<k} col_i u_i / nu_i + u_k
    double* R = s->R.data() + (size_t)k * s->L;
    std::copy(col.begin(), col.end(), R);
    R[k] = n1;
    s->nu[k] = n1;
    s->m = k + 1;
  }
  return kOk;
}

}  // extern "C"
