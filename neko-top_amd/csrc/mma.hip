// mma.hip -- the method of moving asymptotes (Svanberg) with the design
// vectors in HBM.
//
// Replaces neko-top's mma_t (sources/mma/mma.f90:49-458) and its CPU backend
// (sources/mma/bcknd/cpu/mma_cpu.f90): gensub (:8-184), the dual-primal
// interior-point subsolve (:186-579) and the KKT residual (:581-641).  The
// reference runs all of it on the host; here every n-long operation is a
// fused device stream and only the m-sized quantities (y, lambda, s, mu, z,
// zeta, the (m+1)x(m+1) Newton system) live on the host, which is where the
// reference's MPI_Allreduce calls would hand them over anyway.
//
// Per Newton step the device makes four passes over the n variables:
//   newton : delx, diagx and the reductions g_i, bb_i, AA_ik
//   dir    : dx, dxsi, deta, the step-length bound and the line-search copy
//   step   : x, xsi, eta at the trial step and the residual reductions
//            (repeated while the backtracking line search halves the step)
// Reductions are deterministic (fixed trees); sums and maxima come back in one
// read per pass.  The reference's default-real literals (0.01, 0.1, 0.9,
// 1.01 in single precision) are kept.  m is a template parameter (1..4).
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"
#include "nekoamd.h"
#include "reduce.h"

namespace nekoamd {

#define MMA_F0_01 ((double)0.01f)
#define MMA_F0_1 ((double)0.1f)
#define MMA_F0_9 ((double)0.9f)
#define MMA_F1_01 ((double)1.01f)

constexpr int kMmaMaxM = 4;
constexpr int kMB = 256;        // threads per block of an n-pass
constexpr int kMG = 1024;       // max blocks (partials per value)
constexpr int kMaxVals = 2 + 2 * kMmaMaxM + kMmaMaxM * kMmaMaxM;

struct MmaVec {
  int64_t n;
  double *x, *xsi, *eta, *low, *upp, *alpha, *beta, *xmin, *xmax, *p0j, *q0j;
  double *pij, *qij;  // [i*n + j]
  double *xold1, *xold2, *dx, *dxsi, *deta, *xold, *xsiold, *etaold, *delx,
      *diagx;
};

struct MmaState {
  int n = 0, m = 0, max_iter = 100;
  double a0 = 0, asyinit = 0.5, asyincr = 1.2, asydecr = 0.7, epsimin = 0;
  double residumax = HUGE_VAL, residunorm = HUGE_VAL, z = 0, zeta = 0;
  std::vector<double> a, c, d, bi, y, lambda, s, mu;
  MmaVec v{};
  double* mem = nullptr;   // one device block for all vectors
  double* part = nullptr;  // kMaxVals * kMG partials + kMaxVals results
  double* hres = nullptr;  // pinned read-back
  int newton_iters = 0;    // diagnostics of the last update
};

// ---- reductions ------------------------------------------------------------
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Reduce K per-thread values over the block; ismax bit v selects max for
// value v, else sum.  Writes part[v*gridDim.x + blockIdx.x].
template <int K>
__device__ __forceinline__ void block_out(double (&acc)[K], unsigned ismax,
                                          double* __restrict__ part) {
  __shared__ double sh[K][kMB / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int v = 0; v < K; ++v) {
    const double r = (ismax >> v) & 1u ? wave_max(acc[v]) : wave_sum(acc[v]);
    if (lane == 0) sh[v][wv] = r;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    const int v = threadIdx.x;
    double r = sh[v][0];
    for (int q = 1; q < kMB / 64; ++q)
      r = (ismax >> v) & 1u ? fmax(r, sh[v][q]) : r + sh[v][q];
    part[(int64_t)v * gridDim.x + blockIdx.x] = r;
  }
}

__global__ void __launch_bounds__(kFinBlock)
    mma_fin_k(const double* __restrict__ part, int np, unsigned ismax,
              double* __restrict__ out) {
  __shared__ double sh[kFinBlock / 64];
  const int v = blockIdx.x;
  const bool mx = (ismax >> v) & 1u;
  double r = 0.0;
  for (int q = threadIdx.x; q < np; q += kFinBlock) {
    const double p = part[(int64_t)v * np + q];
    r = mx ? fmax(r, p) : r + p;
  }
  r = mx ? wave_max(r) : wave_sum(r);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) sh[wv] = r;
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = sh[0];
    for (int q = 1; q < kFinBlock / 64; ++q) a = mx ? fmax(a, sh[q]) : a + sh[q];
    out[v] = a;
  }
}

// ---- per-variable kernels --------------------------------------------------
template <int M>
__global__ void __launch_bounds__(kMB)
    gensub_k(MmaVec V, const double* __restrict__ xd,
             const double* __restrict__ df0dx, const double* __restrict__ dfdx,
             int early, double asyinit, double asyincr, double asydecr,
             double* __restrict__ part) {
  double acc[M];
#pragma unroll
  for (int i = 0; i < M; ++i) acc[i] = 0.0;
  for (int64_t j = (int64_t)blockIdx.x * kMB + threadIdx.x; j < V.n;
       j += (int64_t)gridDim.x * kMB) {
    const double x = xd[j], xmn = V.xmin[j], xmx = V.xmax[j], w = xmx - xmn;
    double lo, up;
    if (early) {
      lo = x - asyinit * w;
      up = x + asyinit * w;
    } else {
      const double x1 = V.xold1[j], x2 = V.xold2[j];
      lo = V.low[j];
      up = V.upp[j];
      const double osc = (x - x1) * (x1 - x2);
      if (osc < 0) {
        lo = x - asydecr * (x1 - lo);
        up = x + asydecr * (up - x1);
      } else if (osc > 0) {
        lo = x - asyincr * (x1 - lo);
        up = x + asyincr * (up - x1);
      } else {
        lo = x - (x1 - lo);
        up = x + (up - x1);
      }
      lo = fmax(lo, x - 10.0 * w);
      lo = fmin(lo, x - MMA_F0_01 * w);
      up = fmin(up, x + 10.0 * w);
      up = fmax(up, x + MMA_F0_01 * w);
    }
    V.low[j] = lo;
    V.upp[j] = up;
    V.alpha[j] = fmax(fmax(xmn, lo + 0.1 * (x - lo)), x - 0.5 * w);
    V.beta[j] = fmin(fmin(xmx, up - MMA_F0_1 * (up - x)), x + 0.5 * w);
    const double reg = 0.00001 / fmax(0.00001, w);
    const double du = (up - x) * (up - x), dl = (x - lo) * (x - lo);
    const double g0 = df0dx[j];
    V.p0j[j] = du * (1.001 * fmax(g0, 0.0) + 0.001 * fmax(-g0, 0.0) + reg);
    V.q0j[j] = dl * (0.001 * fmax(g0, 0.0) + 1.001 * fmax(-g0, 0.0) + reg);
#pragma unroll
    for (int i = 0; i < M; ++i) {
      const double g = dfdx[j * M + i];
      const double p = du * (1.001 * fmax(g, 0.0) + 0.001 * fmax(-g, 0.0) + reg);
      const double q = dl * (0.001 * fmax(g, 0.0) + 1.001 * fmax(-g, 0.0) + reg);
      V.pij[(int64_t)i * V.n + j] = p;
      V.qij[(int64_t)i * V.n + j] = q;
      acc[i] = acc[i] + p / (up - x) + q / (x - lo);
    }
  }
  block_out<M>(acc, 0u, part);
}

template <int M>
__global__ void __launch_bounds__(kMB) sub_init_k(MmaVec V) {
  for (int64_t j = (int64_t)blockIdx.x * kMB + threadIdx.x; j < V.n;
       j += (int64_t)gridDim.x * kMB) {
    const double x = 0.5 * (V.alpha[j] + V.beta[j]);
    V.x[j] = x;
    V.xsi[j] = fmax(1.0, 1.0 / (x - V.alpha[j]));
    V.eta[j] = fmax(1.0, 1.0 / (V.beta[j] - x));
  }
}

struct Lam {
  double v[kMmaMaxM];
};

// Residual pieces at (x, xsi, eta); with STEP, first x = xold + steg dx etc.
// Values: [sum of rex^2+rexsi^2+reeta^2, max |.|, g_0..g_{M-1}].
template <int M, bool STEP>
__global__ void __launch_bounds__(kMB)
    resid_k(MmaVec V, Lam lam, double epsi, double steg,
            double* __restrict__ part) {
  double acc[2 + M];
#pragma unroll
  for (int q = 0; q < 2 + M; ++q) acc[q] = 0.0;
  for (int64_t j = (int64_t)blockIdx.x * kMB + threadIdx.x; j < V.n;
       j += (int64_t)gridDim.x * kMB) {
    double x, xsi, eta;
    if constexpr (STEP) {
      x = V.xold[j] + steg * V.dx[j];
      xsi = V.xsiold[j] + steg * V.dxsi[j];
      eta = V.etaold[j] + steg * V.deta[j];
      V.x[j] = x;
      V.xsi[j] = xsi;
      V.eta[j] = eta;
    } else {
      x = V.x[j];
      xsi = V.xsi[j];
      eta = V.eta[j];
    }
    const double ux = V.upp[j] - x, xl = x - V.low[j];
    double pl = 0.0, ql = 0.0;
#pragma unroll
    for (int i = 0; i < M; ++i) {
      const double p = V.pij[(int64_t)i * V.n + j], q = V.qij[(int64_t)i * V.n + j];
      pl += p * lam.v[i];
      ql += q * lam.v[i];
      acc[2 + i] = acc[2 + i] + p / ux + q / xl;
    }
    pl = V.p0j[j] + pl;
    ql = V.q0j[j] + ql;
    const double rex = (pl / (ux * ux) - ql / (xl * xl)) - xsi + eta;
    const double rexsi = xsi * (x - V.alpha[j]) - epsi;
    const double reeta = eta * (V.beta[j] - x) - epsi;
    acc[0] += rex * rex + rexsi * rexsi + reeta * reeta;
    acc[1] = fmax(acc[1], fmax(fabs(rex), fmax(fabs(rexsi), fabs(reeta))));
  }
  block_out<2 + M>(acc, 2u, part);
}

// Newton system pieces.  Values: [g_i (M), bb_i (M), AA_ik (M*M)].
template <int M>
__global__ void __launch_bounds__(kMB)
    newton_k(MmaVec V, Lam lam, double epsi, double* __restrict__ part) {
  double acc[2 * M + M * M];
#pragma unroll
  for (int q = 0; q < 2 * M + M * M; ++q) acc[q] = 0.0;
  for (int64_t j = (int64_t)blockIdx.x * kMB + threadIdx.x; j < V.n;
       j += (int64_t)gridDim.x * kMB) {
    const double x = V.x[j], ux = V.upp[j] - x, xl = x - V.low[j];
    const double xa = x - V.alpha[j], bx = V.beta[j] - x;
    double a = 0.0, pl = 0.0, ql = 0.0, GG[M];
#pragma unroll
    for (int i = 0; i < M; ++i) {
      const double p = V.pij[(int64_t)i * V.n + j], q = V.qij[(int64_t)i * V.n + j];
      a = a + p * lam.v[i] / (ux * ux) - q * lam.v[i] / (xl * xl);
      pl += p * lam.v[i];
      ql += q * lam.v[i];
      GG[i] = p / (ux * ux) - q / (xl * xl);
      acc[i] = acc[i] + p / ux + q / xl;
    }
    const double delx = a + V.p0j[j] / (ux * ux) - V.q0j[j] / (xl * xl) -
                        epsi / xa + epsi / bx;
    const double dg = (V.p0j[j] + pl) / (ux * ux * ux) + (V.q0j[j] + ql) / (xl * xl * xl);
    const double diagx = 2.0 * dg + V.xsi[j] / xa + V.eta[j] / bx;
    V.delx[j] = delx;
    V.diagx[j] = diagx;
    const double r = delx / diagx, inv = 1.0 / diagx;
#pragma unroll
    for (int i = 0; i < M; ++i) {
      acc[M + i] = acc[M + i] + GG[i] * r;
#pragma unroll
      for (int k = 0; k < M; ++k)
        acc[2 * M + i * M + k] = acc[2 * M + i * M + k] + GG[i] * inv * GG[k];
    }
  }
  block_out<2 * M + M * M>(acc, 0u, part);
}

// Search direction for x, xsi, eta; saves the line-search origin.
// Value: the largest step-length bound (max).
template <int M>
__global__ void __launch_bounds__(kMB)
    dir_k(MmaVec V, Lam dlam, double epsi, double* __restrict__ part) {
  double acc[1] = {1.0};
  for (int64_t j = (int64_t)blockIdx.x * kMB + threadIdx.x; j < V.n;
       j += (int64_t)gridDim.x * kMB) {
    const double x = V.x[j], ux = V.upp[j] - x, xl = x - V.low[j];
    const double xa = x - V.alpha[j], bx = V.beta[j] - x;
    const double xsi = V.xsi[j], eta = V.eta[j], diagx = V.diagx[j];
    double gl = 0.0;
#pragma unroll
    for (int i = 0; i < M; ++i)
      gl += (V.pij[(int64_t)i * V.n + j] / (ux * ux) -
             V.qij[(int64_t)i * V.n + j] / (xl * xl)) * dlam.v[i];
    const double dx = -V.delx[j] / diagx - gl / diagx;
    const double dxsi = -xsi + (epsi - dx * xsi) / xa;
    const double deta = -eta + (epsi + dx * eta) / bx;
    V.dx[j] = dx;
    V.dxsi[j] = dxsi;
    V.deta[j] = deta;
    V.xold[j] = x;
    V.xsiold[j] = xsi;
    V.etaold[j] = eta;
    double b = fmax(-(MMA_F1_01 * dxsi / xsi), -(MMA_F1_01 * deta / eta));
    b = fmax(b, -(MMA_F1_01 * dx / xa));
    b = fmax(b, MMA_F1_01 * dx / bx);
    acc[0] = fmax(acc[0], b);
  }
  block_out<1>(acc, 1u, part);
}

// KKT pieces at the design x.  Values: [sum of squares, max |.|].
template <int M>
__global__ void __launch_bounds__(kMB)
    kkt_k(MmaVec V, Lam lam, const double* __restrict__ xd,
          const double* __restrict__ df0dx, const double* __restrict__ dfdx,
          double* __restrict__ part) {
  double acc[2] = {0.0, 0.0};
  for (int64_t j = (int64_t)blockIdx.x * kMB + threadIdx.x; j < V.n;
       j += (int64_t)gridDim.x * kMB) {
    double g = 0.0;
#pragma unroll
    for (int i = 0; i < M; ++i) g += dfdx[j * M + i] * lam.v[i];
    const double x = xd[j], xsi = V.xsi[j], eta = V.eta[j];
    const double rex = df0dx[j] + g - xsi + eta;
    const double rexsi = xsi * (x - V.xmin[j]);
    const double reeta = eta * (V.xmax[j] - x);
    acc[0] += rex * rex + rexsi * rexsi + reeta * reeta;
    acc[1] = fmax(acc[1], fmax(fabs(rex), fmax(fabs(rexsi), fabs(reeta))));
  }
  block_out<2>(acc, 2u, part);
}

// ---- host side -------------------------------------------------------------
static int mma_grid(int64_t n) { return nk_grid(n, kMB, kMG); }

static int finish(MmaState* S, int nvals, int grid, unsigned ismax,
                  double* out) {
  NK_LAUNCHED();
  hipLaunchKernelGGL(mma_fin_k, dim3(nvals), dim3(kFinBlock), 0, stream(),
                     S->part, grid, ismax, S->part + kMaxVals * kMG);
  NK_LAUNCHED();
  NK_HIP(hipMemcpyAsync(S->hres, S->part + kMaxVals * kMG,
                        sizeof(double) * nvals, hipMemcpyDeviceToHost,
                        stream()));
  NK_HIP(hipStreamSynchronize(stream()));
  std::memcpy(out, S->hres, sizeof(double) * nvals);
  return kOk;
}

// Gaussian elimination with row partial pivoting (DGESV semantics),
// column-major N x N.  Returns 0 or the 1-based index of a zero pivot.
static int dense_solve(int N, std::vector<double>& A, std::vector<double>& b) {
  int info = 0;
  for (int k = 0; k < N; ++k) {
    int p = k;
    for (int r = k + 1; r < N; ++r)
      if (std::fabs(A[r + k * N]) > std::fabs(A[p + k * N])) p = r;
    if (p != k) {
      for (int c = 0; c < N; ++c) std::swap(A[k + c * N], A[p + c * N]);
      std::swap(b[k], b[p]);
    }
    if (A[k + k * N] == 0.0) {
      if (!info) info = k + 1;
      continue;
    }
    for (int r = k + 1; r < N; ++r) {
      const double l = A[r + k * N] / A[k + k * N];
      A[r + k * N] = l;
      for (int c = k + 1; c < N; ++c) A[r + c * N] -= l * A[k + c * N];
      b[r] -= l * b[k];
    }
  }
  if (info) return info;
  for (int k = N - 1; k >= 0; --k) {
    double s = b[k];
    for (int c = k + 1; c < N; ++c) s -= A[k + c * N] * b[c];
    b[k] = s / A[k + k * N];
  }
  return 0;
}

struct Small {
  double max = 0.0, sq = 0.0;
};

// The m-sized residuals (rey, rez, relambda, remu, rezeta, res).
static Small small_resid(const MmaState* S, const double* g, const double* y,
                         double z, const double* lam, const double* mu,
                         double zeta, const double* s, double epsi) {
  Small r;
  double la = 0.0;
  for (int i = 0; i < S->m; ++i) la += lam[i] * S->a[i];
  const double rez = S->a0 - zeta - la, rezeta = zeta * z - epsi;
  for (int i = 0; i < S->m; ++i) {
    const double rey = S->c[i] + S->d[i] * y[i] - lam[i] - mu[i];
    const double rel = g[i] - S->a[i] * z - y[i] + s[i] - S->bi[i];
    const double remu = mu[i] * y[i] - epsi;
    const double res = lam[i] * s[i] - epsi;
    r.max = std::fmax(r.max, std::fmax(std::fmax(std::fabs(rey), std::fabs(rel)),
                                       std::fmax(std::fabs(remu), std::fabs(res))));
    r.sq += rey * rey + rel * rel + remu * remu + res * res;
  }
  r.max = std::fmax(r.max, std::fmax(std::fabs(rez), std::fabs(rezeta)));
  r.sq += rez * rez + rezeta * rezeta;
  return r;
}

template <int M>
static int mma_update_t(MmaState* S, int iter, double* xd, const double* df0dx,
                        const double* fval, const double* dfdx) {
  MmaVec& V = S->v;
  const int grid = mma_grid(V.n);
  double out[kMaxVals];
  // ---- gensub
  hipLaunchKernelGGL((gensub_k<M>), dim3(grid), dim3(kMB), 0, stream(), V, xd,
                     df0dx, dfdx, iter < 3 ? 1 : 0, S->asyinit, S->asyincr,
                     S->asydecr, S->part);
  int rc = finish(S, M, grid, 0u, out);
  if (rc) return rc;
  for (int i = 0; i < M; ++i) S->bi[i] = out[i] - fval[i];
  // ---- subsolve
  double epsi = 1.0, z = 1.0, zeta = 1.0;
  double y[M], lam[M], s[M], mu[M];
  for (int i = 0; i < M; ++i) {
    y[i] = 1.0;
    lam[i] = 1.0;
    s[i] = 1.0;
    mu[i] = std::fmax(1.0, 0.5 * S->c[i]);
  }
  hipLaunchKernelGGL((sub_init_k<M>), dim3(grid), dim3(kMB), 0, stream(), V);
  NK_LAUNCHED();
  S->newton_iters = 0;
  while (epsi > MMA_F0_9 * S->epsimin) {
    Lam L{};
    for (int i = 0; i < M; ++i) L.v[i] = lam[i];
    hipLaunchKernelGGL((resid_k<M, false>), dim3(grid), dim3(kMB), 0, stream(),
                       V, L, epsi, 0.0, S->part);
    rc = finish(S, 2 + M, grid, 2u, out);
    if (rc) return rc;
    Small sm = small_resid(S, out + 2, y, z, lam, mu, zeta, s, epsi);
    double residumax = std::fmax(out[1], sm.max);
    double residunorm = std::sqrt(sm.sq + out[0]);
    for (int it = 1; it <= S->max_iter; ++it) {
      if (residumax < epsi) break;
      ++S->newton_iters;
      for (int i = 0; i < M; ++i) L.v[i] = lam[i];
      hipLaunchKernelGGL((newton_k<M>), dim3(grid), dim3(kMB), 0, stream(), V,
                         L, epsi, S->part);
      rc = finish(S, 2 * M + M * M, grid, 0u, out);
      if (rc) return rc;
      const double* g = out;
      const double* bbs = out + M;
      const double* AAs = out + 2 * M;
      double la = 0.0;
      for (int i = 0; i < M; ++i) la += lam[i] * S->a[i];
      const double delz = S->a0 - la - epsi / z;
      double dely[M];
      std::vector<double> A((M + 1) * (M + 1), 0.0), bb(M + 1, 0.0);
      for (int i = 0; i < M; ++i) {
        dely[i] = S->c[i] + S->d[i] * y[i] - lam[i] - epsi / y[i];
        const double dellam = g[i] - S->a[i] * z - y[i] - S->bi[i] + epsi / lam[i];
        bb[i] = dellam + dely[i] / (S->d[i] + mu[i] / y[i]) - bbs[i];
      }
      bb[M] = delz;
      for (int i = 0; i < M; ++i)
        for (int k = 0; k < M; ++k) A[i + k * (M + 1)] = AAs[i * M + k];
      for (int i = 0; i < M; ++i) {
        A[i + i * (M + 1)] += s[i] / lam[i] + 1.0 / (S->d[i] + mu[i] / y[i]);
        A[i + M * (M + 1)] = S->a[i];
        A[M + i * (M + 1)] = S->a[i];
      }
      A[M + M * (M + 1)] = -zeta / z;
      const int info = dense_solve(M + 1, A, bb);
      if (info) {
        set_error("mma_update", "the (m+1) Newton system is singular (DGESV)");
        return kErrNumeric;
      }
      double dlam[M], dy[M], dmu[M], ds[M];
      for (int i = 0; i < M; ++i) dlam[i] = bb[i];
      const double dz = bb[M];
      for (int i = 0; i < M; ++i) {
        dy[i] = (-dely[i] + dlam[i]) / (S->d[i] + mu[i] / y[i]);
        dmu[i] = -mu[i] + (epsi - mu[i] * dy[i]) / y[i];
        ds[i] = -s[i] + (epsi - dlam[i] * s[i]) / lam[i];
      }
      const double dzeta = -zeta + (epsi - zeta * dz) / z;
      Lam DL{};
      for (int i = 0; i < M; ++i) DL.v[i] = dlam[i];
      hipLaunchKernelGGL((dir_k<M>), dim3(grid), dim3(kMB), 0, stream(), V, DL,
                         epsi, S->part);
      rc = finish(S, 1, grid, 1u, out);
      if (rc) return rc;
      double steg = out[0];
      for (int i = 0; i < M; ++i) {
        steg = std::fmax(steg, -(MMA_F1_01 * dy[i] / y[i]));
        steg = std::fmax(steg, -(MMA_F1_01 * dlam[i] / lam[i]));
        steg = std::fmax(steg, -(MMA_F1_01 * dmu[i] / mu[i]));
        steg = std::fmax(steg, -(MMA_F1_01 * ds[i] / s[i]));
      }
      steg = std::fmax(steg, -(MMA_F1_01 * dz / z));
      steg = std::fmax(steg, -(MMA_F1_01 * dzeta / zeta));
      steg = 1.0 / steg;
      double yo[M], lo[M], so[M], muo[M];
      std::memcpy(yo, y, sizeof(y));
      std::memcpy(lo, lam, sizeof(lam));
      std::memcpy(so, s, sizeof(s));
      std::memcpy(muo, mu, sizeof(mu));
      const double zo = z, zetao = zeta;
      double newresidu = 2.0 * residunorm, newmax = residumax;
      int itto = 0;
      while (newresidu > residunorm && itto < 50) {
        ++itto;
        for (int i = 0; i < M; ++i) {
          y[i] = yo[i] + steg * dy[i];
          lam[i] = lo[i] + steg * dlam[i];
          mu[i] = muo[i] + steg * dmu[i];
          s[i] = so[i] + steg * ds[i];
        }
        z = zo + steg * dz;
        zeta = zetao + steg * dzeta;
        for (int i = 0; i < M; ++i) L.v[i] = lam[i];
        hipLaunchKernelGGL((resid_k<M, true>), dim3(grid), dim3(kMB), 0,
                           stream(), V, L, epsi, steg, S->part);
        rc = finish(S, 2 + M, grid, 2u, out);
        if (rc) return rc;
        sm = small_resid(S, out + 2, y, z, lam, mu, zeta, s, epsi);
        newresidu = std::sqrt(sm.sq + out[0]);
        newmax = std::fmax(out[1], sm.max);
        steg = steg / 2.0;
      }
      residunorm = newresidu;
      residumax = newmax;
    }
    epsi = 0.1 * epsi;
  }
  // xold2 <- xold1 <- design x ; design x <- x
  NK_HIP(hipMemcpyAsync(V.xold2, V.xold1, sizeof(double) * V.n,
                        hipMemcpyDeviceToDevice, stream()));
  NK_HIP(hipMemcpyAsync(V.xold1, xd, sizeof(double) * V.n,
                        hipMemcpyDeviceToDevice, stream()));
  NK_HIP(hipMemcpyAsync(xd, V.x, sizeof(double) * V.n,
                        hipMemcpyDeviceToDevice, stream()));
  for (int i = 0; i < M; ++i) {
    S->y[i] = y[i];
    S->lambda[i] = lam[i];
    S->s[i] = s[i];
    S->mu[i] = mu[i];
  }
  S->z = z;
  S->zeta = zeta;
  NK_HIP(hipStreamSynchronize(stream()));
  return kOk;
}

template <int M>
static int mma_kkt_t(MmaState* S, const double* xd, const double* df0dx,
                     const double* fval, const double* dfdx) {
  const int grid = mma_grid(S->v.n);
  Lam L{};
  for (int i = 0; i < M; ++i) L.v[i] = S->lambda[i];
  hipLaunchKernelGGL((kkt_k<M>), dim3(grid), dim3(kMB), 0, stream(), S->v, L,
                     xd, df0dx, dfdx, S->part);
  double out[2];
  int rc = finish(S, 2, grid, 2u, out);
  if (rc) return rc;
  double mx = out[1], sq = 0.0, la = 0.0;
  for (int i = 0; i < M; ++i) la += S->lambda[i] * S->a[i];
  const double rez = S->a0 - S->zeta - la, rezeta = S->zeta * S->z;
  for (int i = 0; i < M; ++i) {
    const double rey = S->c[i] + S->d[i] * S->y[i] - S->lambda[i] - S->mu[i];
    const double rel = fval[i] - S->a[i] * S->z - S->y[i] + S->s[i];
    const double remu = S->mu[i] * S->y[i], res = S->lambda[i] * S->s[i];
    mx = std::fmax(mx, std::fmax(std::fmax(std::fabs(rey), std::fabs(rel)),
                                 std::fmax(std::fabs(remu), std::fabs(res))));
    sq += rey * rey + rel * rel + remu * remu + res * res;
  }
  mx = std::fmax(mx, std::fmax(std::fabs(rez), std::fabs(rezeta)));
  sq += rez * rez + rezeta * rezeta;
  S->residumax = mx;
  S->residunorm = std::sqrt(sq + out[0]);
  return kOk;
}

}  // namespace nekoamd

using namespace nekoamd;

extern "C" {

int nekoamd_mma_create(void** mma, const int* n, const int* m, const void* x,
                       const double* a0, const double* a, const double* c,
                       const double* d, const void* xmin, const void* xmax,
                       const int* max_iter, const double* epsimin,
                       const double* asyinit, const double* asyincr,
                       const double* asydecr) {
  NK_INIT();
  NK_REQUIRE(mma && n && m && x && a0 && a && c && d && xmin && xmax &&
                 max_iter && epsimin && asyinit && asyincr && asydecr,
             "null argument");
  NK_REQUIRE(*n > 0, "n must be positive");
  NK_REQUIRE(*m >= 1 && *m <= kMmaMaxM, "m must be in [1, 4]");
  MmaState* S = new MmaState();
  S->n = *n;
  S->m = *m;
  S->max_iter = *max_iter;
  S->a0 = *a0;
  S->asyinit = *asyinit;
  S->asyincr = *asyincr;
  S->asydecr = *asydecr;
  S->epsimin = *epsimin > 0 ? *epsimin : 1.0e-9 * std::sqrt((double)(*m + *n));
  S->a.assign(a, a + *m);
  S->c.assign(c, c + *m);
  S->d.assign(d, d + *m);
  S->bi.assign(*m, 0.0);
  S->y.assign(*m, 0.0);
  S->lambda.assign(*m, 0.0);
  S->s.assign(*m, 0.0);
  S->mu.assign(*m, 0.0);
  const int64_t N = *n;
  const size_t nvec = 21 + 2 * (size_t)*m;
  hipError_t e = hipMalloc((void**)&S->mem, sizeof(double) * nvec * N);
  if (e == hipSuccess)
    e = hipMalloc((void**)&S->part, sizeof(double) * (kMaxVals * kMG + kMaxVals));
  if (e == hipSuccess)
    e = hipHostMalloc((void**)&S->hres, sizeof(double) * kMaxVals,
                      hipHostMallocDefault);
  if (e != hipSuccess) {
    if (S->mem) (void)hipFree(S->mem);
    if (S->part) (void)hipFree(S->part);
    delete S;
    return fail_hip(__func__, e);
  }
  double* p = S->mem;
  MmaVec& V = S->v;
  V.n = N;
  double** slots[] = {&V.x,     &V.xsi,   &V.eta,   &V.low,    &V.upp,
                      &V.alpha, &V.beta,  &V.xmin,  &V.xmax,   &V.p0j,
                      &V.q0j,   &V.xold1, &V.xold2, &V.dx,     &V.dxsi,
                      &V.deta,  &V.xold,  &V.xsiold, &V.etaold, &V.delx,
                      &V.diagx};
  for (double** s : slots) {
    *s = p;
    p += N;
  }
  V.pij = p;
  p += (size_t)*m * N;
  V.qij = p;
  // initial state (mma.f90:132-184): xold1 = xold2 = x, low = min x, upp = max x
  std::vector<double> hx(N);
  e = hipMemcpy(hx.data(), x, sizeof(double) * N, hipMemcpyDeviceToHost);
  if (e == hipSuccess)
    e = hipMemcpy(V.xold1, x, sizeof(double) * N, hipMemcpyDeviceToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(V.xold2, x, sizeof(double) * N, hipMemcpyDeviceToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(V.xmin, xmin, sizeof(double) * N, hipMemcpyDeviceToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(V.xmax, xmax, sizeof(double) * N, hipMemcpyDeviceToDevice);
  if (e != hipSuccess) {
    (void)hipFree(S->mem);
    (void)hipFree(S->part);
    (void)hipHostFree(S->hres);
    delete S;
    return fail_hip(__func__, e);
  }
  double lo = hx[0], hi = hx[0];
  for (int64_t j = 1; j < N; ++j) {
    lo = std::fmin(lo, hx[j]);
    hi = std::fmax(hi, hx[j]);
  }
  int rc = nekoamd_cfill(V.low, &lo, &V.n);
  if (!rc) rc = nekoamd_cfill(V.upp, &hi, &V.n);
  if (rc) {
    (void)hipFree(S->mem);
    (void)hipFree(S->part);
    (void)hipHostFree(S->hres);
    delete S;
    return rc;
  }
  *mma = S;
  return kOk;
}

int nekoamd_mma_free(void* mma) {
  MmaState* S = (MmaState*)mma;
  if (!S) return kOk;
  NK_HIP(hipStreamSynchronize(stream()));
  if (S->mem) NK_HIP(hipFree(S->mem));
  if (S->part) NK_HIP(hipFree(S->part));
  if (S->hres) NK_HIP(hipHostFree(S->hres));
  delete S;
  return kOk;
}

// x (device, n) is replaced by the new design; dfdx is m x n column-major
// on the device (the reference's dfdx(m, n)); fval on the host (m).
int nekoamd_mma_update(void* mma, const int* iter, void* x, const void* df0dx,
                       const double* fval, const void* dfdx) {
  NK_INIT();
  MmaState* S = (MmaState*)mma;
  NK_REQUIRE(S && iter && x && df0dx && fval && dfdx, "null argument");
  switch (S->m) {
    case 1: return mma_update_t<1>(S, *iter, (double*)x, (const double*)df0dx, fval, (const double*)dfdx);
    case 2: return mma_update_t<2>(S, *iter, (double*)x, (const double*)df0dx, fval, (const double*)dfdx);
    case 3: return mma_update_t<3>(S, *iter, (double*)x, (const double*)df0dx, fval, (const double*)dfdx);
    default: return mma_update_t<4>(S, *iter, (double*)x, (const double*)df0dx, fval, (const double*)dfdx);
  }
}

int nekoamd_mma_kkt(void* mma, const void* x, const void* df0dx,
                    const double* fval, const void* dfdx) {
  NK_INIT();
  MmaState* S = (MmaState*)mma;
  NK_REQUIRE(S && x && df0dx && fval && dfdx, "null argument");
  switch (S->m) {
    case 1: return mma_kkt_t<1>(S, (const double*)x, (const double*)df0dx, fval, (const double*)dfdx);
    case 2: return mma_kkt_t<2>(S, (const double*)x, (const double*)df0dx, fval, (const double*)dfdx);
    case 3: return mma_kkt_t<3>(S, (const double*)x, (const double*)df0dx, fval, (const double*)dfdx);
    default: return mma_kkt_t<4>(S, (const double*)x, (const double*)df0dx, fval, (const double*)dfdx);
  }
}

// res[0] = residumax, res[1] = residunorm, res[2] = Newton iterations of
// the last update.
int nekoamd_mma_residuals(void* mma, double* res) {
  MmaState* S = (MmaState*)mma;
  NK_REQUIRE(S && res, "null argument");
  res[0] = S->residumax;
  res[1] = S->residunorm;
  res[2] = S->newton_iters;
  return kOk;
}

// Host copies of the m-sized duals: y, lambda, s, mu (m each), zs = {z, zeta}.
int nekoamd_mma_duals(void* mma, double* y, double* lambda, double* s,
                      double* mu, double* zs) {
  MmaState* S = (MmaState*)mma;
  NK_REQUIRE(S && y && lambda && s && mu && zs, "null argument");
  for (int i = 0; i < S->m; ++i) {
    y[i] = S->y[i];
    lambda[i] = S->lambda[i];
    s[i] = S->s[i];
    mu[i] = S->mu[i];
  }
  zs[0] = S->z;
  zs[1] = S->zeta;
  return kOk;
}

}  // extern "C"
