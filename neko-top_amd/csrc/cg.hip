// cg.hip -- Jacobi preconditioner and preconditioned conjugate gradients
// with device-resident scalars, in FP64 or in FP32 with FP64 reliable
// updates.
//
// Replaces Neko's jacobi_t / device_jacobi_t (precon_factory 'jacobi' at
// sources/adjoint/adjoint_scheme.f90:1068-1076; update() at
// sources/adjoint/adjoint_pnpn.f90:732, :781) and ksp_t%solve for 'cg'
// (krylov_solver_factory at adjoint_scheme.f90:1046-1055; solves at
// adjoint_pnpn.f90:734-736 and, per component, :784-786).
//
// One iteration k is four kernels (plus the gs buckets) on the library
// stream:
//   ax_cg  : z = d r;  p_k = z + beta p_{k-1};  x += alpha_{k-1} p_{k-1};
//            w = A_L p_k;  per-element partials of w.p_k
//   gs     : w = QQ^T w                                         (gs.hip)
//   pap    : 64 first-level sums of the element partials
//   update : alpha = rtz / pap (every block sums the same 64 values in the
//            same order), r = mask(r - alpha w), partials of r.r.mult and
//            r.(d r).mult;  then (one block) beta, ||r||, stop flag
// The x update of iteration k is deferred into iteration k+1's ax_cg (which
// reads p_k anyway) and flushed at the end, so every vector is streamed
// once per iteration: 4 reads + 3 writes in ax_cg besides the operator
// data, 4 reads + 1 write in update.  The Dirichlet mask is folded into a
// copy d_m of the Jacobi diagonal that is zero on masked points: update
// zeroes r there instead of a separate bc pass over w.
//
// alpha, beta, the norm and the stop flag live in device memory, so the host
// never waits inside an iteration; once the stop flag is up every later
// kernel returns immediately (a halt flag, raised with it and, in a mixed
// solve, while a requested reliable update is pending), and the host polls
// the flag every few iterations.  Iteration counts and stopping follow the reference's CG:
// x0 = 0, norm sqrt(<r,r>_mult)/sqrt(volume), checked after each iteration.
//
// The fused pap = sum_local w.p equals the reference's <QQ^T w, p>_mult
// because p is continuous and zero on the Dirichlet mask (r starts as the
// gs-assembled, masked right-hand side and stays so).
//
// Mixed precision (NEKOAMD_CG_MIXED): the iteration vectors r, p, w, d_m,
// mult, h1, h2 are stored in FP32 and x in FP64; every inner product
// accumulates in FP64.  The operator reads FP32 p and writes FP32 w but
// contracts in FP64 registers and LDS: the kernel is bound by HBM, so the
// FP64 arithmetic costs no time, and A p is then exact to FP32 rounding of
// the result instead of to FP32 rounding of terms that cancel (for smooth
// p those terms are ~cond(A) times larger than A p).  The true residual is
// recomputed in FP64 (reliable update: r = f - A x with the FP64 operator,
// r32 = r, search direction kept) whenever the iterated residual has fallen
// by a factor delta below its maximum since the last update, and before the
// solve may stop; the solve stops only on the FP64 residual.  The update
// group is queued after every batch of iterations and is a no-op unless
// requested; once requested, the batch's remaining FP32 iterations are
// no-ops too, so no FP32 iteration runs on a residual that has fallen more
// than delta since the last FP64 one.
#include "ax_core.h"
#include <algorithm>

#include "internal.h"
#include "nekoamd.h"
#include "reduce.h"

namespace nekoamd {

constexpr int kPapParts = 64;       // first-level sums of the element partials
constexpr int kPapBlock = 512;
// Reliable updates: a segment runs until the residual has fallen by delta,
// set from the measured gap between the iterated and the true residual at
// each update (cg_rel_fin_k), within [kMinDelta, kReliableDelta]; the first
// solve of a state starts at kReliableDelta, later ones at the delta their
// predecessor ended with.
constexpr double kReliableDelta = 0.1;
constexpr double kMinDelta = 1e-5;
constexpr double kGapMargin = 100.0;

struct CgScal {
  double rtz1, rtz2, pap, alpha, beta, rtr, rnorm, res_start, norm_fac, abstol;
  double rmax, delta, rseg;
  int iter, max_iter, done, xpend;
  int mixed, no_update, nupdate;
  // the iteration kernels return at once while halt is set: the solve is
  // done, or (mixed) a reliable update was requested and has not yet run
  int halt;
};

static_assert(sizeof(CgScal) % 4 == 0, "CgScal is copied in 4-byte words");

struct CgState {
  int64_t n = 0;
  int nelv = 0;
  int grid = 1;
  int prec = NEKOAMD_CG_FP64;
  CgScal* sc = nullptr;       // device
  double* part_ax = nullptr;  // nelv
  double* part_pap = nullptr; // kPapParts
  double* part = nullptr;     // 3 * kRedGrid
  CgScal* host = nullptr;     // pinned, 2 slots
  // FP64 iteration vectors (w in both modes; r, p, dm in FP64 mode)
  double* r = nullptr;
  double* w = nullptr;
  double* p = nullptr;
  double* dm = nullptr;
  // FP32 iteration vectors (mixed mode)
  float* r32 = nullptr;
  float* p32 = nullptr;
  float* w32 = nullptr;
  float* dm32 = nullptr;
  float* mult32 = nullptr;
  float* h1_32 = nullptr;
  float* h2_32 = nullptr;
  float* dx32 = nullptr;      // mixed: FP32 copy of the derivative matrix
  int last_iter = 0;          // iterations of the previous solve
  // mixed mode: FP32 copies of G11, G12, G13, G22, G23, G33, B (7 n), of
  // mult and of D (dx32), made on the first solve with a given set of those
  // arrays (keyed by their addresses: the geometry of a mesh is fixed once
  // built)
  float* geo32 = nullptr;
  const void* geo_key[9] = {};
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
// Per-solve preparation: out = (TO) in, and the masked diagonal copy.
template <typename TO>
__global__ void __launch_bounds__(256)
    cast_k(int64_t n, TO* __restrict__ out, const double* __restrict__ in) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256)
    out[i] = (TO)in[i];
}

// The per-solve FP32 copies of the mixed solve: Jacobi diagonal, h1, h2.
__global__ void __launch_bounds__(256)
    cast3_k(int64_t n, float* __restrict__ o1, const double* __restrict__ i1,
            float* __restrict__ o2, const double* __restrict__ i2,
            float* __restrict__ o3, const double* __restrict__ i3) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    o1[i] = (float)i1[i];
    o2[i] = (float)i2[i];
    if (o3) o3[i] = (float)i3[i];
  }
}

template <typename T>
__global__ void __launch_bounds__(256)
    zero_at_k(T* __restrict__ u, const int32_t* __restrict__ mask,
              int64_t nmask) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < nmask;
       i += (int64_t)gridDim.x * 256)
    u[mask[i]] = T(0);
}

// x = 0, r = f (and r64 = f when r64 is given: the FP64 solves' r is r64
// itself, the mixed solves carry r32 only); partials of f.f.mult and
// r.(d r).mult.
template <typename T>
__global__ void __launch_bounds__(kRedBlock)
    cg_start_k(int64_t n, double* __restrict__ x, T* __restrict__ r,
               const double* __restrict__ f, const T* __restrict__ dm,
               const T* __restrict__ mult, double* __restrict__ r64,
               double* __restrict__ part) {
  __shared__ double sh[kRedBlock / 64];
  double a = 0.0, b = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRedBlock) {
    const double fi = f[i];
    const T ri = (T)fi;
    x[i] = 0.0;
    r[i] = ri;
    if (r64 != nullptr) r64[i] = fi;
    const double m = (double)mult[i];
    a += m * fi * fi;
    b += m * (double)ri * (double)(ri * dm[i]);
  }
  a = block_sum<kRedBlock>(a, sh);
  b = block_sum<kRedBlock>(b, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[kRedGrid + blockIdx.x] = b;
  }
}

// cg_start_k<float> for the three components of a coupled mixed solve in
// one pass (dm and mult read once): per component the same operations and
// block partials.
struct Cg3Start {
  double* x[3];
  float* r[3];
  const double* f[3];
  double* part[3];
};

__global__ void __launch_bounds__(kRedBlock)
    cg_start3_k(int64_t n, Cg3Start v, const float* __restrict__ dm,
                const float* __restrict__ mult) {
  __shared__ double sh[kRedBlock / 64];
  double a[3] = {0.0, 0.0, 0.0}, b[3] = {0.0, 0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRedBlock) {
    const double m = (double)mult[i];
    const float d = dm[i];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double fi = v.f[c][i];
      const float ri = (float)fi;
      v.x[c][i] = 0.0;
      v.r[c][i] = ri;
      a[c] += m * fi * fi;
      b[c] += m * (double)ri * (double)(ri * d);
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double sa = block_sum<kRedBlock>(a[c], sh);
    const double sb = block_sum<kRedBlock>(b[c], sh);
    if (threadIdx.x == 0) {
      v.part[c][blockIdx.x] = sa;
      v.part[c][kRedGrid + blockIdx.x] = sb;
    }
  }
}

__global__ void __launch_bounds__(kFinBlock)
    cg_start_fin_k(CgScal* __restrict__ sc, const double* __restrict__ part,
                   int np, double volume, double abstol, int max_iter,
                   int mixed) {
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
    sc->rmax = rn;
    sc->rseg = rn;
    // a state's previous solve set delta from the gap it measured; the
    // systems of one state differ from solve to solve only in their
    // right-hand side and coefficients, so its next solve starts from it
    if (!(sc->delta >= kMinDelta && sc->delta <= kReliableDelta))
      sc->delta = kReliableDelta;
    sc->rtz1 = rtz;
    sc->rtz2 = 1.0;
    sc->beta = 0.0;
    sc->alpha = 0.0;
    sc->pap = 0.0;
    sc->iter = 0;
    sc->max_iter = max_iter;
    sc->xpend = 0;
    sc->mixed = mixed;
    sc->no_update = 1;
    sc->nupdate = 0;
    sc->done = (rn == 0.0 || max_iter <= 0) ? 1 : 0;
    sc->halt = sc->done;
  }
}

// ---------------------------------------------------------------------------
// Fused search-direction update, deferred x update and operator, one
// thread per GLL point (ax_core.h).  p, w, r, d_m, h1, h2 are stored as T,
// x as FP64; the contraction runs in FP64.  (x in FP32 between reliable
// updates, as a correction folded into the FP64 x at each update, saves 8 B
// per point but its rounding is a rough field whose image under A is up to
// cond(A) times the residual it tracks; the iterated residual then drifts
// from the true one within a segment and the pressure solves took up to
// 2.7x the iterations.)  TA: the contraction type -- FP64, or FP32 for the
// mixed Helmholtz (velocity) systems, which are mass-dominated
// (h2 = rho bd0/dt + chi against h1 = mu), so the FP32 rounding of the
// stiffness terms that cancel is small against A p; the pressure systems
// contract in FP64 (in FP32 their first 32^3 solve took 2573 iterations
// against 2120).
template <int LX, typename T, typename TG, bool H2, typename TA = double>
__global__ void __launch_bounds__(LX * LX * LX)
    ax_cg_k(T* __restrict__ p, double* __restrict__ x, T* __restrict__ w,
            const T* __restrict__ r, const T* __restrict__ dm,
            const double* __restrict__ dx, const T* __restrict__ h1,
            const T* __restrict__ h2, const TG* __restrict__ g11,
            const TG* __restrict__ g12, const TG* __restrict__ g13,
            const TG* __restrict__ g22, const TG* __restrict__ g23,
            const TG* __restrict__ g33, const TG* __restrict__ bm,
            const CgScal* __restrict__ sc, double* __restrict__ part_ax) {
  constexpr int NS = LX * LX, NP = NS * LX;
  __shared__ TA su[NP];
  __shared__ TA sr[NP];
  __shared__ TA ss[NP];
  __shared__ TA st[NP];
  __shared__ TA sD[NS];
  __shared__ double sd[NP % 64 == 0 ? NP / 64 : NP];
  if (sc->halt) return;
  const bool first = sc->iter == 0;
  const bool xpend = sc->xpend != 0;
  const T beta = (T)sc->beta;
  const double alpha = sc->alpha;
  const int t = threadIdx.x;
  const int64_t g = (int64_t)blockIdx.x * NP + t;
  if (t < NS) sD[t] = (TA)dx[t];
  const T z = r[g] * dm[g];
  T pn = z;
  if (!first) {
    const T po = p[g];
    pn = z + beta * po;
    if (xpend)
      __builtin_nontemporal_store(
          __builtin_nontemporal_load(x + g) + alpha * (double)po, x + g);
  }
  p[g] = pn;
  const AxPoint<TA> a =
      ax_load<TA, T, TG, H2>(g, h1, h2, g11, g12, g13, g22, g23, g33, bm);
  const TA pd = (TA)pn;
  su[t] = pd;
  __syncthreads();
  const T wt = (T)ax_point<LX, TA, H2>(pd, a, su, sr, ss, st, sD, t);
  w[g] = wt;
  const double dot = element_sum<NP>((double)wt * (double)pn, sd);
  if (t == 0) part_ax[blockIdx.x] = dot;
}

// The three velocity components of a coupled solve in one launch (lx = 8):
// the operator data (G, h1, h2, B) are read once per element and applied to
// each active component in turn, with exactly ax_cg_k's arithmetic per
// component.  A component is skipped when the driver left it out (active 0)
// or its solve has halted on the device (a uniform branch: every thread of
// the workgroup reads the same flag).
struct Cg3Vec {
  float* p[3];
  double* x[3];
  float* w[3];
  const float* r[3];
  const CgScal* sc[3];
  double* part[3];
  int active[3];
};

template <int LX, typename TG, bool H2, typename TA>
__global__ void __launch_bounds__(LX * LX * LX)
    ax_cg3_k(Cg3Vec v, const float* __restrict__ dm,
             const double* __restrict__ dx, const TA* __restrict__ dxa,
             const float* __restrict__ h1,
             const float* __restrict__ h2, const TG* __restrict__ g11,
             const TG* __restrict__ g12, const TG* __restrict__ g13,
             const TG* __restrict__ g22, const TG* __restrict__ g23,
             const TG* __restrict__ g33, const TG* __restrict__ bm) {
  typedef float T;
  constexpr int NS = LX * LX, NP = NS * LX;
  __shared__ TA su[NP];
  __shared__ TA sr[NP];
  __shared__ TA ss[NP];
  __shared__ TA st[NP];
  __shared__ TA sD[NS];
  __shared__ double sd[NP % 64 == 0 ? NP / 64 : NP];
  const int t = threadIdx.x;
  const int64_t g = (int64_t)blockIdx.x * NP + t;
  if (t < NS) sD[t] = (TA)dx[t];
  const AxPoint<TA> a =
      ax_load<TA, T, TG, H2>(g, h1, h2, g11, g12, g13, g22, g23, g33, bm);
  const T d = dm[g];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (!v.active[c]) continue;
    const CgScal* sc = v.sc[c];
    if (sc->halt) continue;
    const bool first = sc->iter == 0;
    const bool xpend = sc->xpend != 0;
    const T beta = (T)sc->beta;
    const double alpha = sc->alpha;
    const T z = v.r[c][g] * d;
    T pn = z;
    if (!first) {
      const T po = v.p[c][g];
      pn = z + beta * po;
      if (xpend)
        __builtin_nontemporal_store(
            __builtin_nontemporal_load(v.x[c] + g) + alpha * (double)po,
            v.x[c] + g);
    }
    v.p[c][g] = pn;
    const TA pd = (TA)pn;
    su[t] = pd;
    __syncthreads();
    T wt;
    if constexpr (LX * LX == 64)
      wt = (T)ax_point_sk<LX, TA, H2>(pd, a, su, sr, ss, st, sD, dxa, t);
    else
      wt = (T)ax_point<LX, TA, H2>(pd, a, su, sr, ss, st, sD, t);
    v.w[c][g] = wt;
    // its barrier also orders this component's LDS reads before the next
    // component's writes
    const double dot = element_sum<NP>((double)wt * (double)pn, sd);
    if (t == 0) v.part[c][blockIdx.x] = dot;
  }
}

// Small elements (lx = 2, 3: the coarse level of the multigrid): EPB
// elements per workgroup, one thread per point, so a 64-wide wave is full
// (one workgroup of 8 threads per element would leave 56 lanes idle and
// launch nelv workgroups).  Per element exactly the arithmetic of ax_cg_k
// (ax_point on the element's slice of the LDS tiles, the serial element
// sum), so the results are the same.
template <int LX, int EPB, typename T, typename TG, bool H2>
__global__ void __launch_bounds__(LX * LX * LX * EPB)
    ax_cg_mk(T* __restrict__ p, double* __restrict__ x, T* __restrict__ w,
             const T* __restrict__ r, const T* __restrict__ dm,
             const double* __restrict__ dx, const T* __restrict__ h1,
             const T* __restrict__ h2, const TG* __restrict__ g11,
             const TG* __restrict__ g12, const TG* __restrict__ g13,
             const TG* __restrict__ g22, const TG* __restrict__ g23,
             const TG* __restrict__ g33, const TG* __restrict__ bm,
             const CgScal* __restrict__ sc, int nelv,
             double* __restrict__ part_ax) {
  constexpr int NS = LX * LX, NP = NS * LX, NB = NP * EPB;
  __shared__ double su[NB];
  __shared__ double sr[NB];
  __shared__ double ss[NB];
  __shared__ double st[NB];
  __shared__ double sD[NS];
  __shared__ double sd[NB];
  if (sc->halt) return;
  const bool first = sc->iter == 0;
  const bool xpend = sc->xpend != 0;
  const T beta = (T)sc->beta;
  const double alpha = sc->alpha;
  const int el = threadIdx.x / NP, t = threadIdx.x % NP;
  const int e = blockIdx.x * EPB + el;
  const bool live = e < nelv;
  const int64_t g = (int64_t)(live ? e : 0) * NP + t;
  if (threadIdx.x < NS) sD[threadIdx.x] = dx[threadIdx.x];
  T pn = T(0);
  AxPoint<double> a{};
  if (live) {
    const T z = r[g] * dm[g];
    pn = z;
    if (!first) {
      const T po = p[g];
      pn = z + beta * po;
      if (xpend) x[g] = x[g] + alpha * (double)po;
    }
    p[g] = pn;
    a = ax_load<double, T, TG, H2>(g, h1, h2, g11, g12, g13, g22, g23, g33,
                                   bm);
  }
  const double pd = (double)pn;
  const int o = el * NP;
  su[o + t] = pd;
  __syncthreads();
  const T wt =
      (T)ax_point<LX, double, H2>(pd, a, su + o, sr + o, ss + o, st + o, sD, t);
  sd[o + t] = (double)wt * pd;
  if (live) w[g] = wt;
  __syncthreads();
  if (live && t == 0) {
    double s = 0.0;
    for (int q = 0; q < NP; ++q) s += sd[o + q];
    part_ax[e] = s;
  }
}

// First-level sums of the element partials: block b sums the contiguous
// chunk b of part_ax.
__device__ __forceinline__ void pap_body(const double* __restrict__ part_ax,
                                         int nelv,
                                         double* __restrict__ part_pap,
                                         const CgScal* __restrict__ sc) {
  __shared__ double sh[kPapBlock / 64];
  if (sc->halt) return;
  const int chunk = (nelv + kPapParts - 1) / kPapParts;
  const int lo = blockIdx.x * chunk;
  const int hi = min(lo + chunk, nelv);
  double s = 0.0;
  for (int q = lo + (int)threadIdx.x; q < hi; q += kPapBlock) s += part_ax[q];
  s = block_sum<kPapBlock>(s, sh);
  if (threadIdx.x == 0) part_pap[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kPapBlock)
    cg_pap_k(const double* __restrict__ part_ax, int nelv,
             double* __restrict__ part_pap, const CgScal* __restrict__ sc) {
  pap_body(part_ax, nelv, part_pap, sc);
}

template <typename T>
__global__ void __launch_bounds__(kRedBlock)
    cg_update_k(int64_t n, T* __restrict__ r, const T* __restrict__ w,
                const T* __restrict__ dm, const T* __restrict__ mult,
                CgScal* __restrict__ sc, const double* __restrict__ part_pap,
                double* __restrict__ part) {
  __shared__ double sh[kRedBlock / 64];
  __shared__ double s_alpha;
  if (sc->halt) return;
  if (threadIdx.x < 64) {
    // kPapParts == 64: one value per lane, the same butterfly in every block
    const double pap = wave_sum(part_pap[threadIdx.x]);
    if (threadIdx.x == 0) {
      const double al = sc->rtz1 / pap;
      s_alpha = al;
      if (blockIdx.x == 0) {
        sc->pap = pap;
        sc->alpha = al;
      }
    }
  }
  __syncthreads();
  const T alpha = (T)s_alpha;
  double a = 0.0, b = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRedBlock) {
    const T di = dm[i];
    const T ri = di == T(0) ? T(0) : r[i] - alpha * w[i];
    r[i] = ri;
    const double rm = (double)ri * (double)mult[i];
    a += rm * (double)ri;
    b += rm * (double)(ri * di);
  }
  a = block_sum<kRedBlock>(a, sh);
  b = block_sum<kRedBlock>(b, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[kRedGrid + blockIdx.x] = b;
  }
}

__device__ __forceinline__ void update_fin_body(CgScal* __restrict__ sc,
                                                const double* __restrict__ part,
                                                int np) {
  if (sc->halt) return;
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
    sc->xpend = 1;
    // a non-finite residual ends the solve (reported as not converged)
    const bool stop = rn < sc->abstol || it >= sc->max_iter || !isfinite(rn);
    if (sc->mixed) {
      // A requested reliable update holds the FP32 iterations until it has
      // run: iterating on past it lets the iterated residual fall below
      // FP32's resolution of the true one, after which the search
      // directions no longer follow the true residual.
      sc->rmax = fmax(sc->rmax, rn);
      if (stop || rn < sc->delta * sc->rmax) {
        sc->no_update = 0;
        sc->halt = 1;
      }
      if (stop) sc->done = 1;
    } else if (stop) {
      sc->done = 1;
      sc->halt = 1;
    }
  }
}

__global__ void __launch_bounds__(kFinBlock)
    cg_update_fin_k(CgScal* __restrict__ sc, const double* __restrict__ part,
                    int np) {
  update_fin_body(sc, part, np);
}

// The reductions and residual update of the three components of a coupled
// FP32 solve in one launch each: component c is block row blockIdx.y of the
// pap and fin kernels (the same bodies as cg_pap_k and cg_update_fin_k), and
// cg_update3_k streams dm and mult once for all three, with cg_update_k's
// arithmetic and sums per component.  Components left out (active 0) or
// halted on the device are skipped.
struct Cg3Red {
  const double* part_ax[3];
  double* part_pap[3];
  double* part[3];
  CgScal* sc[3];
  float* r[3];
  const float* w[3];
  int active[3];
};

__global__ void __launch_bounds__(kPapBlock) cg_pap3_k(Cg3Red v, int nelv) {
  const int c = blockIdx.y;
  if (!v.active[c]) return;
  pap_body(v.part_ax[c], nelv, v.part_pap[c], v.sc[c]);
}

__global__ void __launch_bounds__(kFinBlock) cg_update_fin3_k(Cg3Red v, int np) {
  const int c = blockIdx.y;
  if (!v.active[c]) return;
  update_fin_body(v.sc[c], v.part[c], np);
}

__global__ void __launch_bounds__(kRedBlock)
    cg_update3_k(int64_t n, Cg3Red v, const float* __restrict__ dm,
                 const float* __restrict__ mult) {
  __shared__ double sh[kRedBlock / 64];
  __shared__ double s_alpha[3];
  bool act[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) act[c] = v.active[c] && !v.sc[c]->halt;
  if (threadIdx.x < 64) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (!act[c]) continue;
      // kPapParts == 64: one value per lane, the same butterfly in every block
      const double pap = wave_sum(v.part_pap[c][threadIdx.x]);
      if (threadIdx.x == 0) {
        const double al = v.sc[c]->rtz1 / pap;
        s_alpha[c] = al;
        if (blockIdx.x == 0) {
          v.sc[c]->pap = pap;
          v.sc[c]->alpha = al;
        }
      }
    }
  }
  __syncthreads();
  float alpha[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) alpha[c] = act[c] ? (float)s_alpha[c] : 0.0f;
  double a[3] = {0.0, 0.0, 0.0}, b[3] = {0.0, 0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRedBlock) {
    const float di = dm[i];
    const double mi = (double)mult[i];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (!act[c]) continue;
      const float ri = di == 0.0f ? 0.0f : v.r[c][i] - alpha[c] * v.w[c][i];
      v.r[c][i] = ri;
      const double rm = (double)ri * mi;
      a[c] += rm * (double)ri;
      b[c] += rm * (double)(ri * di);
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (!act[c]) continue;
    const double sa = block_sum<kRedBlock>(a[c], sh);
    const double sb = block_sum<kRedBlock>(b[c], sh);
    if (threadIdx.x == 0) {
      v.part[c][blockIdx.x] = sa;
      v.part[c][kRedGrid + blockIdx.x] = sb;
    }
  }
}

// x += alpha p for the pending last direction: at the end of an FP64
// solve (gate == null), and as the first step of a reliable update (gate =
// &no_update: a no-op unless an update was requested).
template <typename T>
__global__ void __launch_bounds__(256)
    cg_flush_k(int64_t n, double* __restrict__ x, const T* __restrict__ p,
               const CgScal* __restrict__ sc, const int* __restrict__ gate) {
  if (gate != nullptr && *gate) return;
  if (!sc->xpend) return;
  const double alpha = sc->alpha;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256)
    x[i] = x[i] + alpha * (double)p[i];
}

// Reliable update, mixed mode; every kernel is a no-op unless requested.

// r64 = mask(f - w64) with w64 = QQ^T A x64 (FP64), r32 = r64; partials of
// r64.r64.mult, r32.(d r32).mult and of the squared gap between the true
// residual and the iterated one it replaces, (r64 - r32_old)^2 mult.  r64
// itself is not stored: the mixed iteration carries r32 only.
__global__ void __launch_bounds__(kRedBlock)
    cg_rel_r_k(int64_t n, float* __restrict__ r32,
               const double* __restrict__ f, const double* __restrict__ w64,
               const float* __restrict__ dm, const float* __restrict__ mult,
               const CgScal* __restrict__ sc, double* __restrict__ part) {
  __shared__ double sh[kRedBlock / 64];
  if (sc->no_update) return;
  double a = 0.0, b = 0.0, c = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRedBlock) {
    const float di = dm[i];
    const double ri = di == 0.0f ? 0.0 : f[i] - w64[i];
    const double gi = ri - (double)r32[i];
    const float rf = (float)ri;
    r32[i] = rf;
    const double m = (double)mult[i];
    a += m * ri * ri;
    b += m * (double)rf * (double)(rf * di);
    c += m * gi * gi;
  }
  a = block_sum<kRedBlock>(a, sh);
  b = block_sum<kRedBlock>(b, sh);
  c = block_sum<kRedBlock>(c, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[kRedGrid + blockIdx.x] = b;
    part[2 * kRedGrid + blockIdx.x] = c;
  }
}

__global__ void __launch_bounds__(kFinBlock)
    cg_rel_fin_k(CgScal* __restrict__ sc, const double* __restrict__ part,
                 int np) {
  if (sc->no_update) return;
  __shared__ double sh[kFinBlock / 64];
  const double rtr = sum_partials<kFinBlock>(part, np, sh);
  const double rtz = sum_partials<kFinBlock>(part + kRedGrid, np, sh);
  const double gap2 = sum_partials<kFinBlock>(part + 2 * kRedGrid, np, sh);
  __syncthreads();
  if (threadIdx.x == 0) {
    const double rn = sqrt(rtr) * sc->norm_fac;
    // The gap between the iterated and the true residual grows in
    // proportion to the true residual at the segment's start, rseg (FP32
    // rounding of r, p and the x correction, and the FP32 operator data).
    // The next segment may run until the residual has fallen by
    // delta = kGapMargin gap / rseg, so that the gap stays at about
    // 1/kGapMargin of the residual when the next update replaces it.
    const double gap = sqrt(gap2) * sc->norm_fac;
    if (sc->rseg > 0.0)
      sc->delta = fmin(kReliableDelta,
                       fmax(kMinDelta, kGapMargin * gap / sc->rseg));
    sc->rseg = rn;
    sc->rtr = rtr;
    sc->rnorm = rn;
    sc->rmax = rn;
    sc->beta = rtz / sc->rtz2;
    sc->rtz1 = rtz;
    sc->xpend = 0;
    sc->nupdate = sc->nupdate + 1;
    sc->done = (rn < sc->abstol || sc->iter >= sc->max_iter || !isfinite(rn)) ? 1 : 0;
    sc->no_update = 1;
    sc->halt = sc->done;
  }
}

// ---------------------------------------------------------------------------
template <typename T, typename TG>
static int ax_cg_launch(int lx, int nelv, bool ifh2, T* p, double* x, T* w,
                        const T* r, const T* dm, const double* dx, const T* h1,
                        const T* h2, const TG* g11, const TG* g12,
                        const TG* g13, const TG* g22, const TG* g23,
                        const TG* g33, const TG* B, const CgScal* sc,
                        double* part_ax) {
  if (nelv == 0) return kOk;
#define NK_AXCG_CASE(L)                                                       \
  case L:                                                                     \
    if (ifh2)                                                                 \
      nk_launch(ax_cg_k<L, T, TG, true, T>, dim3(nelv), dim3(L * L * L), 0,   \
                p, x, w, r, dm, dx, h1, h2, g11, g12, g13, g22, g23, g33, B,  \
                sc, part_ax);                                                 \
    else                                                                      \
      nk_launch(ax_cg_k<L, T, TG, false>, dim3(nelv), dim3(L * L * L), 0, p,  \
                x, w, r, dm, dx, h1, h2, g11, g12, g13, g22, g23, g33, B, sc, \
                part_ax);                                                     \
    break;
#define NK_AXCG_MCASE(L, E)                                                   \
  case L: {                                                                   \
    const dim3 grid((nelv + E - 1) / E), blk(L * L * L * E);                  \
    if (ifh2)                                                                 \
      nk_launch(ax_cg_mk<L, E, T, TG, true>, grid, blk, 0, p, x, w, r, dm,    \
                dx, h1, h2, g11, g12, g13, g22, g23, g33, B, sc, nelv,        \
                part_ax);                                                     \
    else                                                                      \
      nk_launch(ax_cg_mk<L, E, T, TG, false>, grid, blk, 0, p, x, w, r, dm,   \
                dx, h1, h2, g11, g12, g13, g22, g23, g33, B, sc, nelv,        \
                part_ax);                                                     \
    break;                                                                    \
  }
  switch (lx) {
    NK_AXCG_MCASE(2, 32)
    NK_AXCG_MCASE(3, 8)
    NK_AXCG_CASE(4)
    NK_AXCG_CASE(5)
    NK_AXCG_CASE(6)
    NK_AXCG_CASE(7)
    NK_AXCG_CASE(8)
    NK_AXCG_CASE(9)
    NK_AXCG_CASE(10)
    default:
      return fail_arg("cg", "lx must be in [2, 10]");
  }
#undef NK_AXCG_CASE
#undef NK_AXCG_MCASE
  NK_LAUNCHED();
  return kOk;
}

static CgState* as_cg(void* cg) { return (CgState*)cg; }

// The operator and space a solve runs on.
struct CgOp {
  void* gs;
  const int32_t* mask;
  int64_t nmask;
  const double* dx;
  const double *h1, *h2, *g11, *g12, *g13, *g22, *g23, *g33, *B;
  int lx;
  bool ifh2;
};

// The geometric factors and mass an iteration reads, in storage type TG.
template <typename TG>
struct Geo {
  const TG *g11, *g12, *g13, *g22, *g23, *g33, *B;
};

static Geo<double> geo64(const CgOp& op) {
  return {op.g11, op.g12, op.g13, op.g22, op.g23, op.g33, op.B};
}

template <typename T, typename TG>
static int cg_iteration(CgState* s, const CgOp& op, const Geo<TG>& G,
                        double* x, T* r, T* p, T* w, const T* dm,
                        const T* mult, const T* h1, const T* h2) {
  int rc = ax_cg_launch<T, TG>(op.lx, s->nelv, op.ifh2, p, x, w, r, dm, op.dx,
                               h1, h2, G.g11, G.g12, G.g13, G.g22, G.g23,
                               G.g33, G.B, s->sc, s->part_ax);
  if (rc) return rc;
  rc = gs_launch<T>(op.gs, w, NEKOAMD_GS_ADD, &s->sc->halt);
  if (rc) return rc;
  nk_launch(cg_pap_k, dim3(kPapParts), dim3(kPapBlock), 0, s->part_ax,
            s->nelv, s->part_pap, s->sc);
  NK_LAUNCHED();
  nk_launch(cg_update_k<T>, dim3(s->grid), dim3(kRedBlock), 0, s->n, r, w, dm,
            mult, s->sc, s->part_pap, s->part);
  NK_LAUNCHED();
  nk_launch(cg_update_fin_k, dim3(1), dim3(kFinBlock), 0, s->sc, s->part,
            s->grid);
  NK_LAUNCHED();
  return kOk;
}

static int cg_reliable(CgState* s, const CgOp& op, double* x64,
                       const double* f, const float* dm32 = nullptr,
                       const float* mult32 = nullptr) {
  if (!dm32) dm32 = s->dm32;
  if (!mult32) mult32 = s->mult32;
  const int g = nk_grid(s->n, 256);
  hipLaunchKernelGGL(cg_flush_k<float>, dim3(g), dim3(256), 0, stream(), s->n,
                     x64, s->p32, s->sc, &s->sc->no_update);
  NK_LAUNCHED();
  int rc = ax_helm_launch<double, double>(
      s->w, x64, op.dx, op.h1, op.h2, op.g11, op.g12, op.g13, op.g22, op.g23,
      op.g33, op.B, op.lx, s->nelv, op.ifh2, &s->sc->no_update);
  if (rc) return rc;
  rc = gs_launch<double>(op.gs, s->w, NEKOAMD_GS_ADD, &s->sc->no_update);
  if (rc) return rc;
  hipLaunchKernelGGL(cg_rel_r_k, dim3(s->grid), dim3(kRedBlock), 0, stream(),
                     s->n, s->r32, f, s->w, dm32, mult32, s->sc, s->part);
  NK_LAUNCHED();
  hipLaunchKernelGGL(cg_rel_fin_k, dim3(1), dim3(kFinBlock), 0, stream(),
                     s->sc, s->part, s->grid);
  NK_LAUNCHED();
  return kOk;
}

// The reliable updates of the components of a coupled solve in `active`:
// cg_reliable's kernels per component, the FP64 operator of all of them in
// one launch (ax_helm3_launch; each component's result is the same as its
// own launch's).
static int cg_reliable3(CgState* const* s, const CgOp& op, double* const* X,
                        const double* const* F, const float* dm32,
                        const float* mult32, const bool* active) {
  double* W[3];
  const double* U[3];
  const int* skip[3];
  for (int c = 0; c < 3; ++c) {
    W[c] = s[c]->w;
    U[c] = X[c];
    skip[c] = &s[c]->sc->no_update;
    if (!active[c]) continue;
    hipLaunchKernelGGL(cg_flush_k<float>, dim3(nk_grid(s[c]->n, 256)),
                       dim3(256), 0, stream(), s[c]->n, X[c], s[c]->p32,
                       s[c]->sc, &s[c]->sc->no_update);
    NK_LAUNCHED();
  }
  int rc = ax_helm3_launch(W, U, skip, active, op.dx, op.h1, op.h2, op.g11,
                           op.g12, op.g13, op.g22, op.g23, op.g33, op.B,
                           op.lx, s[0]->nelv, op.ifh2);
  if (rc) return rc;
  rc = gs_launch3<double>(op.gs, W, NEKOAMD_GS_ADD, skip, active);
  if (rc) return rc;
  for (int c = 0; c < 3; ++c) {
    if (!active[c]) continue;
    CgState* sc = s[c];
    hipLaunchKernelGGL(cg_rel_r_k, dim3(sc->grid), dim3(kRedBlock), 0,
                       stream(), sc->n, sc->r32, F[c], sc->w, dm32, mult32,
                       sc->sc, sc->part);
    NK_LAUNCHED();
    hipLaunchKernelGGL(cg_rel_fin_k, dim3(1), dim3(kFinBlock), 0, stream(),
                       sc->sc, sc->part, sc->grid);
    NK_LAUNCHED();
  }
  return kOk;
}

// Mixed mode: the FP32 geometry, mass and mult (once per set of arrays) and
// this solve's FP32 Jacobi diagonal (masked), h1, h2.
static int prepare_mixed(CgState* s, const CgOp& op, const double* d,
                         const double* mult, Geo<float>& G32) {
  const int64_t n = s->n;
  const int ge = nk_grid(n, 256);
  const void* key[9] = {op.g11, op.g12, op.g13, op.g22, op.g23,
                        op.g33, op.B,   mult,   op.dx};
  if (!std::equal(key, key + 9, s->geo_key)) {
    if (!s->geo32) NK_HIP(hipMalloc((void**)&s->geo32, sizeof(float) * 7 * n));
    const double* src[7] = {op.g11, op.g12, op.g13, op.g22,
                            op.g23, op.g33, op.B};
    for (int q = 0; q < 7; ++q) {
      if (!src[q]) continue;  // B and the cross terms may be absent
      hipLaunchKernelGGL(cast_k<float>, dim3(ge), dim3(256), 0, stream(), n,
                         s->geo32 + (int64_t)q * n, src[q]);
      NK_LAUNCHED();
    }
    hipLaunchKernelGGL(cast_k<float>, dim3(ge), dim3(256), 0, stream(), n,
                       s->mult32, mult);
    NK_LAUNCHED();
    if (!s->dx32) NK_HIP(hipMalloc((void**)&s->dx32, sizeof(float) * 16 * 16));
    hipLaunchKernelGGL(cast_k<float>, dim3(1), dim3(256), 0, stream(),
                       (int64_t)op.lx * op.lx, s->dx32, op.dx);
    NK_LAUNCHED();
    std::copy(key, key + 9, s->geo_key);
  }
  G32 = Geo<float>{s->geo32,         s->geo32 + n,     s->geo32 + 2 * n,
                   s->geo32 + 3 * n, s->geo32 + 4 * n, s->geo32 + 5 * n,
                   s->geo32 + 6 * n};
  hipLaunchKernelGGL(cast3_k, dim3(ge), dim3(256), 0, stream(), n, s->dm32, d,
                     s->h1_32, op.h1, op.ifh2 ? s->h2_32 : nullptr, op.h2);
  NK_LAUNCHED();
  if (op.nmask > 0) {
    hipLaunchKernelGGL(zero_at_k<float>, dim3(nk_grid(op.nmask, 256)),
                       dim3(256), 0, stream(), s->dm32, op.mask, op.nmask);
    NK_LAUNCHED();
  }
  return kOk;
}

// One FP32 iteration of the three mixed solves of a coupled solve: per
// component in `active` the kernels of cg_iteration on the operator data
// the three share (a component that has converged on the device runs them
// as no-ops; the driver leaves out the ones it knows to be done).  At lx = 8
// (the velocity form) the operator of all three components is one launch
// (ax_cg3_k), which reads the operator data once: 322 us for three
// components against 3 x 120-200 us at 32^3 p=7.  (With the contraction in
// FP64 such a launch had measured slower: three FP64 contractions per
// workgroup cost more LDS time than the operator data they save; in FP32
// they cost half.)
static int cg_iteration3(CgState* const* s, const CgOp& op,
                         const Geo<float>& G, double* const* X, const float* dm,
                         const float* mult, const float* h1, const float* h2,
                         const bool* active) {
  if (op.lx == 8 && op.ifh2 && s[0]->nelv > 0) {
    Cg3Vec v{};
    for (int c = 0; c < 3; ++c) {
      v.p[c] = s[c]->p32;
      v.x[c] = X[c];
      v.w[c] = s[c]->w32;
      v.r[c] = s[c]->r32;
      v.sc[c] = s[c]->sc;
      v.part[c] = s[c]->part_ax;
      v.active[c] = active[c] ? 1 : 0;
    }
    nk_launch(ax_cg3_k<8, float, true, float>, dim3(s[0]->nelv), dim3(512), 0,
              v, dm, op.dx, (const float*)s[0]->dx32, h1, h2, G.g11, G.g12,
              G.g13, G.g22, G.g23, G.g33, G.B);
    NK_LAUNCHED();
    float* W[3];
    const int* halt[3];
    Cg3Red red{};
    for (int c = 0; c < 3; ++c) {
      W[c] = s[c]->w32;
      halt[c] = &s[c]->sc->halt;
      red.part_ax[c] = s[c]->part_ax;
      red.part_pap[c] = s[c]->part_pap;
      red.part[c] = s[c]->part;
      red.sc[c] = s[c]->sc;
      red.r[c] = s[c]->r32;
      red.w[c] = s[c]->w32;
      red.active[c] = active[c] ? 1 : 0;
    }
    int rc = gs_launch3<float>(op.gs, W, NEKOAMD_GS_ADD, halt, active);
    if (rc) return rc;
    nk_launch(cg_pap3_k, dim3(kPapParts, 3), dim3(kPapBlock), 0, red,
              s[0]->nelv);
    NK_LAUNCHED();
    nk_launch(cg_update3_k, dim3(s[0]->grid), dim3(kRedBlock), 0, s[0]->n, red,
              dm, mult);
    NK_LAUNCHED();
    nk_launch(cg_update_fin3_k, dim3(1, 3), dim3(kFinBlock), 0, red,
              s[0]->grid);
    NK_LAUNCHED();
    return kOk;
  }
  for (int c = 0; c < 3; ++c) {
    if (!active[c]) continue;
    const int rc = cg_iteration<float>(s[c], op, G, X[c], s[c]->r32,
                                       s[c]->p32, s[c]->w32, dm, mult, h1, h2);
    if (rc) return rc;
  }
  return kOk;
}

static int cg_read(CgState* s, int slot, hipEvent_t ev) {
  if (const int rc = to_host(&s->host[slot], s->sc, sizeof(CgScal)))
    return rc;
  NK_HIP(hipEventRecord(ev, stream()));
  return kOk;
}

static void fill_monitor(const CgScal& h, nekoamd_cg_monitor_t* mon) {
  mon->iter = h.iter;
  mon->done = h.done;
  mon->res_start = h.res_start;
  mon->res_final = h.rnorm;
  mon->nupdate = h.nupdate;
  mon->pad = 0;
}

// Queue batches of `batch` iterations (and, mixed, a reliable-update group
// after each) until the device reports done.  Double-buffered polling: batch
// b+1 is queued before the status read after batch b is waited for, so the
// device never idles on the host.  Every decision reads a copy of a status
// read whose event has completed (`last`), never a slot a copy may still be
// writing.
// In mixed mode a batch can end early (the iterated residual met the
// tolerance but the FP64 one did not, or a reliable update held the rest of
// the batch), so batches are queued until the device itself reports done or
// max_iter iterations.
template <typename F>
static int cg_drive(CgState* s, int max_iter, int batch, F&& queue_batch) {
  hipEvent_t ev[2];
  NK_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  hipError_t e1 = hipEventCreateWithFlags(&ev[1], hipEventDisableTiming);
  if (e1 != hipSuccess) {
    (void)hipEventDestroy(ev[0]);
    return fail_hip("cg_solve", e1);
  }
  int cur = 0, pending = -1, queued = 0, status = kOk;
  const bool mixed = s->prec == NEKOAMD_CG_MIXED;
  CgScal last{};
  bool have = false;  // last holds a completed read
  auto wait = [&](int slot) {
    const hipError_t e = hipEventSynchronize(ev[slot]);
    if (e != hipSuccess) {
      status = fail_hip("cg_solve", e);
      return false;
    }
    last = s->host[slot];
    have = true;
    return true;
  };
  // Short solves (the previous one took fewer than `batch` iterations): a
  // first batch of the previous count + 1, waited on before anything else
  // is queued, so a solve like its predecessor queues at most one iteration
  // past convergence instead of up to two batches; if it falls short, the
  // rest goes in double-buffered batches of 2.
  bool finished = false;
  if (s->last_iter > 0 && s->last_iter < batch) {
    const int nb = s->last_iter + 1;
    status = queue_batch(nb);
    queued += nb;
    if (!status) status = cg_read(s, cur, ev[cur]);
    finished = status != kOk || !wait(cur) || last.done;
    cur ^= 1;
    batch = 2;
  }
  while (!finished) {
    int launched = -1;
    const bool more = queued < max_iter ||
                      (mixed && have && !last.done && last.iter < max_iter);
    if (more) {
      status = queue_batch(batch);
      queued += batch;
      if (!status) status = cg_read(s, cur, ev[cur]);
      if (status) break;
      launched = cur;
      cur ^= 1;
    }
    if (pending >= 0 && (!wait(pending) || last.done)) break;
    if (launched < 0) break;
    pending = launched;
  }
  (void)hipStreamSynchronize(stream());
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  return status;
}

// cg_drive for the three states of a coupled solve: each status read copies
// the three device structs; the solve ends when all three are done (a
// converged component's kernels skip on its done flag meanwhile).
template <typename F>
static int cg_drive3(CgState* const* s, int max_iter, int batch,
                     F&& queue_batch) {
  hipEvent_t ev[2];
  NK_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  hipError_t e1 = hipEventCreateWithFlags(&ev[1], hipEventDisableTiming);
  if (e1 != hipSuccess) {
    (void)hipEventDestroy(ev[0]);
    return fail_hip("cg_solve3", e1);
  }
  int status = kOk;
  auto read3 = [&](int slot) -> int {
    HostCopy hc{};
    hc.n = 3;
    for (int c = 0; c < 3; ++c) {
      hc.dst[c] = &s[c]->host[slot];
      hc.src[c] = s[c]->sc;
      hc.words[c] = (int)(sizeof(CgScal) / 4);
    }
    if (const int rc = to_host(hc)) return rc;
    NK_HIP(hipEventRecord(ev[slot], stream()));
    return kOk;
  };
  // the newest completed read of each component; a component seen done
  // there is no longer queued
  CgScal last[3] = {};
  bool have = false;
  auto wait = [&](int slot) {
    const hipError_t e = hipEventSynchronize(ev[slot]);
    if (e != hipSuccess) {
      status = fail_hip("cg_solve3", e);
      return false;
    }
    for (int c = 0; c < 3; ++c) last[c] = s[c]->host[slot];
    have = true;
    return true;
  };
  auto all_done = [&]() {
    return have && last[0].done && last[1].done && last[2].done;
  };
  auto any_short = [&]() {
    if (!have) return false;
    for (int c = 0; c < 3; ++c)
      if (!last[c].done && last[c].iter < max_iter) return true;
    return false;
  };
  int cur = 0, pending = -1, queued = 0;
  bool finished = false;
  const int lmax = std::max({s[0]->last_iter, s[1]->last_iter, s[2]->last_iter});
  int lim[3];
  if (lmax > 0 && lmax < batch) {
    // first batch: each component's count of the previous solve, plus one
    for (int c = 0; c < 3; ++c)
      lim[c] = (s[c]->last_iter > 0 ? s[c]->last_iter : lmax) + 1;
    const int nb = lmax + 1;
    status = queue_batch(nb, lim);
    queued += nb;
    if (!status) status = read3(cur);
    finished = status != kOk || !wait(cur) || all_done();
    cur ^= 1;
    batch = 2;
  }
  while (!finished) {
    int launched = -1;
    const bool more = queued < max_iter || any_short();
    if (more) {
      for (int c = 0; c < 3; ++c) lim[c] = have && last[c].done ? 0 : batch;
      status = queue_batch(batch, lim);
      queued += batch;
      if (!status) status = read3(cur);
      if (status) break;
      launched = cur;
      cur ^= 1;
    }
    if (pending >= 0 && (!wait(pending) || all_done())) break;
    if (launched < 0) break;
    pending = launched;
  }
  (void)hipStreamSynchronize(stream());
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  return status;
}

template <typename T>
static T* alloc_vec(int64_t n, hipError_t& e) {
  T* p = nullptr;
  if (e == hipSuccess && n > 0) e = hipMalloc((void**)&p, sizeof(T) * n);
  return p;
}

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

int nekoamd_cg_create(void** cg, const int64_t* n, const int* nelv,
                      const int* precision) {
  NK_INIT();
  NK_REQUIRE(cg && n && nelv && precision && *n >= 0 && *nelv >= 0,
             "bad arguments");
  NK_REQUIRE(*precision == NEKOAMD_CG_FP64 || *precision == NEKOAMD_CG_MIXED,
             "precision must be NEKOAMD_CG_FP64 or NEKOAMD_CG_MIXED");
  NK_REQUIRE(*nelv == 0 || *n % *nelv == 0, "n is not a multiple of nelv");
  CgState* s = new CgState();
  s->n = *n;
  s->nelv = *nelv;
  s->prec = *precision;
  s->grid = nk_grid(*n, kRedBlock, kRedGrid);
  const int64_t N = s->n;
  hipError_t e = hipMalloc((void**)&s->sc, sizeof(CgScal));
  if (e == hipSuccess)
    e = hipMalloc((void**)&s->part_ax, sizeof(double) * (s->nelv + 1));
  if (e == hipSuccess)
    e = hipMalloc((void**)&s->part_pap, sizeof(double) * kPapParts);
  if (e == hipSuccess)
    e = hipMalloc((void**)&s->part, sizeof(double) * 3 * kRedGrid);
  if (e == hipSuccess)
    e = hipHostMalloc((void**)&s->host, 2 * sizeof(CgScal),
                      hipHostMallocDefault);
  s->w = alloc_vec<double>(N, e);
  if (s->prec == NEKOAMD_CG_FP64) {
    s->r = alloc_vec<double>(N, e);
    s->p = alloc_vec<double>(N, e);
    s->dm = alloc_vec<double>(N, e);
  } else {
    s->r32 = alloc_vec<float>(N, e);
    s->p32 = alloc_vec<float>(N, e);
    s->w32 = alloc_vec<float>(N, e);
    s->dm32 = alloc_vec<float>(N, e);
    s->mult32 = alloc_vec<float>(N, e);
    s->h1_32 = alloc_vec<float>(N, e);
    s->h2_32 = alloc_vec<float>(N, e);
  }
  if (e != hipSuccess) {
    nekoamd_cg_free(s);
    return fail_hip(__func__, e);
  }
  // a fresh state counts as converged until a solve starts
  CgScal init{};
  init.done = 1;
  init.halt = 1;
  init.no_update = 1;
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
  void* dev[] = {s->sc,    s->part_ax, s->part_pap, s->part,  s->r,
                 s->w,     s->p,       s->dm,       s->r32,   s->p32,
                 s->w32,   s->dm32,    s->mult32,   s->h1_32, s->h2_32,
                 s->geo32, s->dx32};
  for (void* q : dev)
    if (q) NK_HIP(hipFree(q));
  if (s->host) NK_HIP(hipHostFree(s->host));
  delete s;
  return kOk;
}

int nekoamd_cg_status(void* cg, nekoamd_cg_monitor_t* mon) {
  NK_INIT();
  CgState* s = as_cg(cg);
  NK_REQUIRE(s && mon, "null argument");
  if (const int rc = to_host(&s->host[0], s->sc, sizeof(CgScal))) return rc;
  NK_HIP(hipStreamSynchronize(stream()));
  fill_monitor(s->host[0], mon);
  return kOk;
}

int nekoamd_cg_solve(void* cg, void* x, const void* f, const void* d,
                     const void* mult, const double* volume,
                     const double* abstol, const int* max_iter,
                     const int* poll, void* gs, const void* mask,
                     const int64_t* nmask, const void* dx, const void* h1,
                     const void* h2, const void* g11, const void* g12,
                     const void* g13, const void* g22, const void* g23,
                     const void* g33, const void* B, const int* lx,
                     const int* ifh2, nekoamd_cg_monitor_t* mon) {
  NK_INIT();
  CgState* s = as_cg(cg);
  NK_REQUIRE(s && volume && abstol && max_iter && poll && nmask && lx &&
                 ifh2 && mon,
             "null argument");
  NK_REQUIRE(*volume > 0.0, "volume must be positive");
  NK_REQUIRE(*nmask >= 0 && (*nmask == 0 || mask), "bad mask");
  NK_REQUIRE(*lx >= 2 && *lx <= NEKOAMD_LX_MAX, "lx must be in [2, 10]");
  NK_REQUIRE((int64_t)s->nelv * (*lx) * (*lx) * (*lx) == s->n,
             "lx does not match the CG state size");
  NK_REQUIRE(s->n == 0 || (x && f && d && mult && gs && dx && h1 && g11 &&
                           g12 && g13 && g22 && g23 && g33),
             "null field");
  NK_REQUIRE(!*ifh2 || (h2 && B), "ifh2 set but h2 or B is null");
  const int batch = *poll < 1 ? 1 : *poll;
  const int64_t n = s->n;
  const int ge = nk_grid(n, 256);
  const int gm = nk_grid(*nmask, 256);
  CgOp op{gs,
          (const int32_t*)mask,
          *nmask,
          (const double*)dx,
          (const double*)h1,
          (const double*)h2,
          (const double*)g11,
          (const double*)g12,
          (const double*)g13,
          (const double*)g22,
          (const double*)g23,
          (const double*)g33,
          (const double*)B,
          *lx,
          *ifh2 != 0};
  const double* F = (const double*)f;
  int status = kOk;
  if (s->prec == NEKOAMD_CG_FP64) {
    if (n > 0)
      NK_HIP(hipMemcpyAsync(s->dm, d, sizeof(double) * n,
                            hipMemcpyDeviceToDevice, stream()));
    if (*nmask > 0) {
      hipLaunchKernelGGL(zero_at_k<double>, dim3(gm), dim3(256), 0, stream(),
                         s->dm, op.mask, *nmask);
      NK_LAUNCHED();
    }
    double* X = (double*)x;
    hipLaunchKernelGGL(cg_start_k<double>, dim3(s->grid), dim3(kRedBlock), 0,
                       stream(), n, X, s->r, F, s->dm, (const double*)mult,
                       nullptr, s->part);
    NK_LAUNCHED();
    hipLaunchKernelGGL(cg_start_fin_k, dim3(1), dim3(kFinBlock), 0, stream(),
                       s->sc, s->part, s->grid, *volume, *abstol, *max_iter,
                       0);
    NK_LAUNCHED();
    status = cg_drive(s, *max_iter, batch, [&](int nb) {
      for (int q = 0; q < nb; ++q) {
        int rc = cg_iteration<double>(s, op, geo64(op), X, s->r, s->p, s->w, s->dm,
                                      (const double*)mult, op.h1, op.h2);
        if (rc) return rc;
      }
      return (int)kOk;
    });
    if (status) return status;
    hipLaunchKernelGGL(cg_flush_k<double>, dim3(ge), dim3(256), 0, stream(),
                       n, X, s->p, s->sc, nullptr);
    NK_LAUNCHED();
  } else {
    Geo<float> G32;
    int rc0 = prepare_mixed(s, op, (const double*)d, (const double*)mult, G32);
    if (rc0) return rc0;
    double* X = (double*)x;
    hipLaunchKernelGGL(cg_start_k<float>, dim3(s->grid), dim3(kRedBlock), 0,
                       stream(), n, X, s->r32, F, s->dm32, s->mult32,
                       (double*)nullptr, s->part);
    NK_LAUNCHED();
    hipLaunchKernelGGL(cg_start_fin_k, dim3(1), dim3(kFinBlock), 0, stream(),
                       s->sc, s->part, s->grid, *volume, *abstol, *max_iter,
                       1);
    NK_LAUNCHED();
    status = cg_drive(s, *max_iter, batch, [&](int nb) {
      for (int q = 0; q < nb; ++q) {
        int rc = cg_iteration<float>(s, op, G32, X, s->r32, s->p32, s->w32,
                                     s->dm32, s->mult32, s->h1_32, s->h2_32);
        if (rc) return rc;
      }
      return cg_reliable(s, op, X, F);
    });
    if (status) return status;
  }
  const int rc = nekoamd_cg_status(cg, mon);
  if (rc == kOk) s->last_iter = mon->iter;
  return rc;
}

int nekoamd_cg_solve3(void* const* cg, void* const* x, const void* const* f,
                      const void* d, const void* mult, const double* volume,
                      const double* abstol, const int* max_iter,
                      const int* poll, void* gs, const void* mask,
                      const int64_t* nmask, const void* dx, const void* h1,
                      const void* h2, const void* g11, const void* g12,
                      const void* g13, const void* g22, const void* g23,
                      const void* g33, const void* B, const int* lx,
                      const int* ifh2, nekoamd_cg_monitor_t* mon) {
  NK_INIT();
  NK_REQUIRE(cg && x && f && mon, "null argument");
  CgState* s[3] = {as_cg(cg[0]), as_cg(cg[1]), as_cg(cg[2])};
  NK_REQUIRE(s[0] && s[1] && s[2], "null CG state");
  NK_REQUIRE(s[0] != s[1] && s[0] != s[2] && s[1] != s[2],
             "the three CG states must be distinct");
  for (int c = 1; c < 3; ++c)
    NK_REQUIRE(s[c]->n == s[0]->n && s[c]->nelv == s[0]->nelv &&
                   s[c]->prec == s[0]->prec,
               "the three CG states differ in size or precision");
  NK_REQUIRE(lx && volume && abstol && max_iter && poll && nmask && ifh2,
             "null argument");
  // FP64 states, or an element size without the three-direction kernel:
  // three solves one after the other
  if (s[0]->prec != NEKOAMD_CG_MIXED || *lx < 4 || *lx > 10) {
    for (int c = 0; c < 3; ++c) {
      const int rc = nekoamd_cg_solve(
          cg[c], x[c], f[c], d, mult, volume, abstol, max_iter, poll, gs, mask,
          nmask, dx, h1, h2, g11, g12, g13, g22, g23, g33, B, lx, ifh2,
          &mon[c]);
      if (rc) return rc;
    }
    return kOk;
  }
  NK_REQUIRE(*volume > 0.0, "volume must be positive");
  NK_REQUIRE(*nmask >= 0 && (*nmask == 0 || mask), "bad mask");
  NK_REQUIRE((int64_t)s[0]->nelv * (*lx) * (*lx) * (*lx) == s[0]->n,
             "lx does not match the CG state size");
  const int64_t n = s[0]->n;
  NK_REQUIRE(n == 0 || (x[0] && x[1] && x[2] && f[0] && f[1] && f[2] && d &&
                        mult && gs && dx && h1 && g11 && g12 && g13 && g22 &&
                        g23 && g33),
             "null field");
  NK_REQUIRE(!*ifh2 || (h2 && B), "ifh2 set but h2 or B is null");
  const int batch = *poll < 1 ? 1 : *poll;
  CgOp op{gs,
          (const int32_t*)mask,
          *nmask,
          (const double*)dx,
          (const double*)h1,
          (const double*)h2,
          (const double*)g11,
          (const double*)g12,
          (const double*)g13,
          (const double*)g22,
          (const double*)g23,
          (const double*)g33,
          (const double*)B,
          *lx,
          *ifh2 != 0};
  // the operator's FP32 data live in the first state and serve all three
  CgState* s0 = s[0];
  Geo<float> G32;
  int rc = prepare_mixed(s0, op, (const double*)d, (const double*)mult, G32);
  if (rc) return rc;
  double* X[3] = {(double*)x[0], (double*)x[1], (double*)x[2]};
  const double* F[3] = {(const double*)f[0], (const double*)f[1],
                        (const double*)f[2]};
  Cg3Start st{};
  for (int c = 0; c < 3; ++c) {
    st.x[c] = X[c];
    st.r[c] = s[c]->r32;
    st.f[c] = F[c];
    st.part[c] = s[c]->part;
  }
  hipLaunchKernelGGL(cg_start3_k, dim3(s0->grid), dim3(kRedBlock), 0, stream(),
                     n, st, (const float*)s0->dm32, (const float*)s0->mult32);
  NK_LAUNCHED();
  for (int c = 0; c < 3; ++c) {
    hipLaunchKernelGGL(cg_start_fin_k, dim3(1), dim3(kFinBlock), 0, stream(),
                       s[c]->sc, s[c]->part, s[c]->grid, *volume, *abstol,
                       *max_iter, 1);
    NK_LAUNCHED();
  }
  // nb iterations, component c in the first lim[c] of them (lim[c] = 0: c
  // is known to be done; its reliable update is left out too)
  rc = cg_drive3(s, *max_iter, batch, [&](int nb, const int* lim) {
    for (int q = 0; q < nb; ++q) {
      const bool act[3] = {q < lim[0], q < lim[1], q < lim[2]};
      int r = cg_iteration3(s, op, G32, X, s0->dm32, s0->mult32, s0->h1_32,
                            s0->h2_32, act);
      if (r) return r;
    }
    const bool upd[3] = {lim[0] > 0, lim[1] > 0, lim[2] > 0};
    return cg_reliable3(s, op, X, F, s0->dm32, s0->mult32, upd);
  });
  if (rc) return rc;
  for (int c = 0; c < 3; ++c) {
    rc = nekoamd_cg_status(cg[c], &mon[c]);
    if (rc) return rc;
    s[c]->last_iter = mon[c].iter;
  }
  return kOk;
}

}  // extern "C"
