// dealias.hip -- dealiased (over-integrated) advection: the velocity fields
// are interpolated from the lx GLL points to lxd = 3 lx / 2 Gauss-Legendre
// (GL) points per direction, the convective products are formed and
// weighted there, and the result returns to GLL through the transpose of the
// interpolation (a weak-form projection).  The three operators of
// sources/adjoint/adv_adjoint_dealias.f90 and Neko's dealiased forward term:
//
//   forward    f_c = beta f_c - J^T [ t . opgrad(t_c) ]               (Neko
//              adv_dealias_t%compute)
//   linearised f_c = beta f_c - J^T [ u' . opgrad(U_c) + U . opgrad(u'_c) ]
//              (compute_linear_advection_dealias :466-653)
//   adjoint    f_c = beta f_c - J^T [ sum_d a_d opgrad_c(U_d)
//                                     + sum_b D_b^T (w3 a_c sum_d U_d c_b[d]) ]
//              (compute_adjoint_advection_dealias :227-449)
//
// with t = (J x J x J) u on the GL points, opgrad_a(f) = w3 sum_b c_b[a]
// d_b f, w3 the GL tensor weights and c_b[a] the GL metric: the GLL cofactors
// interpolated to GL, as init_dealias builds coef_GL (:155-163).
//
// One workgroup per element does everything on chip: the GLL fields are
// staged in LDS, interpolated to GL by three tensor passes, combined point
// by point in registers (PPT points per thread), and mapped back by three
// transposed passes straight into f.  The reference's device branch makes
// ~20 round trips through n_GL-sized HBM arrays per call; here HBM sees the
// GLL inputs, the GL metric and f once.  The work is FP64 tensor
// contractions of length lx and lxd, so the kernel is bounded by FP64 issue,
// not HBM (about 1.6 MFLOP per element for the adjoint at lx = 8).
#include <cmath>
#include <vector>

#include "internal.h"
#include "nekoamd.h"

namespace nekoamd {
namespace {

constexpr int kFwd = 0, kLin = 1, kAdj = 2;

template <int LXD>
constexpr int dl_threads() {
  return LXD == 12 ? 512 : 256;
}

struct DealiasState {
  int lx = 0, lxd = 0, nelv = 0;
  int64_t n = 0, ngl = 0;
  double* cof = nullptr;   // [9][ngl], row m = 3b + a holds c_b[a]
  double* mats = nullptr;  // J [lxd*lx] | Dg [lxd*lxd] | wg [lxd]
};

// LDS of one element: the 1-D matrices, the two interpolation temporaries
// (t2 also stages one GLL input field) and NB GL-point fields.
template <int LX, int LXD, int NB>
struct DlShared {
  double J[LXD * LX];
  double Dg[LXD * LXD];
  double w[LXD];
  double t1[LXD * LX * LX];
  double t2[LXD * LXD * LX];
  double buf[NB][LXD * LXD * LXD];
};

template <int LX, int LXD, int NB>
__device__ __forceinline__ void load_mats(DlShared<LX, LXD, NB>& sh,
                                          const double* __restrict__ mats) {
  constexpr int NT = dl_threads<LXD>();
  for (int i = threadIdx.x; i < LXD * LX; i += NT) sh.J[i] = mats[i];
  for (int i = threadIdx.x; i < LXD * LXD; i += NT)
    sh.Dg[i] = mats[LXD * LX + i];
  for (int i = threadIdx.x; i < LXD; i += NT)
    sh.w[i] = mats[LXD * LX + LXD * LXD + i];
}

// dst (GL, LDS) = (J x J x J) sh.t2[0 .. LX^3), the staged GLL field.  The
// caller synchronised after staging; returns synchronised.
template <int LX, int LXD, int NB>
__device__ __forceinline__ void interp_up(DlShared<LX, LXD, NB>& sh,
                                          double* dst) {
  constexpr int NT = dl_threads<LXD>();
  constexpr int NPG = LXD * LXD * LXD;
  for (int p = threadIdx.x; p < LXD * LX * LX; p += NT) {
    const int q = p % LXD, jk = p / LXD;
    double s = 0.0;
#pragma unroll
    for (int l = 0; l < LX; ++l) s += sh.J[q * LX + l] * sh.t2[l + LX * jk];
    sh.t1[p] = s;
  }
  __syncthreads();
  for (int p = threadIdx.x; p < LXD * LXD * LX; p += NT) {
    const int q = p % LXD, r = (p / LXD) % LXD, k = p / (LXD * LXD);
    double s = 0.0;
#pragma unroll
    for (int l = 0; l < LX; ++l)
      s += sh.J[r * LX + l] * sh.t1[q + LXD * (l + LX * k)];
    sh.t2[p] = s;
  }
  __syncthreads();
  for (int p = threadIdx.x; p < NPG; p += NT) {
    const int qr = p % (LXD * LXD), tt = p / (LXD * LXD);
    double s = 0.0;
#pragma unroll
    for (int l = 0; l < LX; ++l)
      s += sh.J[tt * LX + l] * sh.t2[qr + LXD * LXD * l];
    dst[p] = s;
  }
  __syncthreads();
}

// Stage the element's GLL values of src and interpolate them to dst.
template <int LX, int LXD, int NB>
__device__ __forceinline__ void load_up(DlShared<LX, LXD, NB>& sh,
                                        const double* __restrict__ src,
                                        int64_t eb, double* dst) {
  constexpr int NT = dl_threads<LXD>();
  for (int p = threadIdx.x; p < LX * LX * LX; p += NT)
    sh.t2[p] = __builtin_nontemporal_load(src + eb + p);
  __syncthreads();
  interp_up(sh, dst);
}

// f[eb + .] = beta f - (J x J x J)^T src, src a GL field in LDS.  The caller
// synchronised after writing src; returns synchronised.
template <int LX, int LXD, int NB>
__device__ __forceinline__ void down_store(DlShared<LX, LXD, NB>& sh,
                                           const double* src,
                                           double* __restrict__ f, int64_t eb,
                                           double beta) {
  constexpr int NT = dl_threads<LXD>();
  for (int p = threadIdx.x; p < LXD * LXD * LX; p += NT) {
    const int qr = p % (LXD * LXD), k = p / (LXD * LXD);
    double s = 0.0;
#pragma unroll
    for (int l = 0; l < LXD; ++l)
      s += sh.J[l * LX + k] * src[qr + LXD * LXD * l];
    sh.t2[p] = s;
  }
  __syncthreads();
  for (int p = threadIdx.x; p < LXD * LX * LX; p += NT) {
    const int q = p % LXD, j = (p / LXD) % LX, k = p / (LXD * LX);
    double s = 0.0;
#pragma unroll
    for (int l = 0; l < LXD; ++l)
      s += sh.J[l * LX + j] * sh.t2[q + LXD * (l + LXD * k)];
    sh.t1[p] = s;
  }
  __syncthreads();
  for (int p = threadIdx.x; p < LX * LX * LX; p += NT) {
    const int i = p % LX, jk = p / LX;
    double s = 0.0;
#pragma unroll
    for (int l = 0; l < LXD; ++l) s += sh.J[l * LX + i] * sh.t1[l + LXD * jk];
    const int64_t g = eb + p;
    f[g] = (beta == 0.0 ? 0.0 : beta * f[g]) - s;
  }
  __syncthreads();
}

// r/s/t derivatives of the GL field f at (i, j, k).
template <int LXD>
__device__ __forceinline__ void gl_grad(const double* Dg, const double* f,
                                        int i, int j, int k, double& r,
                                        double& s, double& t) {
  constexpr int NS = LXD * LXD;
  r = 0.0;
  s = 0.0;
  t = 0.0;
  // unrolled by 2 only: full unrolling lets the scheduler hoist the LDS
  // reads of all points at once and spill
#pragma unroll 2
  for (int m = 0; m < LXD; ++m) {
    r += Dg[i * LXD + m] * f[k * NS + j * LXD + m];
    s += Dg[j * LXD + m] * f[k * NS + m * LXD + i];
    t += Dg[k * LXD + m] * f[m * NS + j * LXD + i];
  }
}

template <int MODE>
constexpr int dl_nbuf() {
  return MODE == kFwd ? 3 : MODE == kLin ? 6 : 4;
}

// a*: forward u,v,w; linearised u'; adjoint the adjoint velocity.
// b*: the base flow U (linearised and adjoint).
template <int LX, int LXD, int MODE>
__global__ void __launch_bounds__(dl_threads<LXD>())
    dealias_k(double* __restrict__ fx, double* __restrict__ fy,
              double* __restrict__ fz, const double* __restrict__ a0,
              const double* __restrict__ a1, const double* __restrict__ a2,
              const double* __restrict__ b0, const double* __restrict__ b1,
              const double* __restrict__ b2, const double* __restrict__ cof,
              const double* __restrict__ mats, int64_t ngl, double beta) {
  constexpr int NT = dl_threads<LXD>();
  constexpr int NP = LX * LX * LX, NPG = LXD * LXD * LXD;
  constexpr int PPT = (NPG + NT - 1) / NT;
  constexpr int NB = dl_nbuf<MODE>();
  __shared__ DlShared<LX, LXD, NB> sh;
  const int e = blockIdx.x;
  const int64_t eb = (int64_t)e * NP, gb = (int64_t)e * NPG;
  double* const f[3] = {fx, fy, fz};
  load_mats(sh, mats);

  if constexpr (MODE == kFwd) {
    load_up(sh, a0, eb, sh.buf[0]);
    load_up(sh, a1, eb, sh.buf[1]);
    load_up(sh, a2, eb, sh.buf[2]);
  } else {
    load_up(sh, b0, eb, sh.buf[0]);
    load_up(sh, b1, eb, sh.buf[1]);
    load_up(sh, b2, eb, sh.buf[2]);
  }
  if constexpr (MODE == kLin) {
    load_up(sh, a0, eb, sh.buf[3]);
    load_up(sh, a1, eb, sh.buf[4]);
    load_up(sh, a2, eb, sh.buf[5]);
  }
  double A[3][PPT];  // adjoint: the adjoint velocity at the thread's points
  if constexpr (MODE == kAdj) {
    const double* src[3] = {a0, a1, a2};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      load_up(sh, src[d], eb, sh.buf[3]);
#pragma unroll
      for (int q = 0; q < PPT; ++q) {
        const int p = threadIdx.x + q * NT;
        A[d][q] = p < NPG ? sh.buf[3][p] : 0.0;
      }
    }
  }

  double O[3][PPT];              // the per-component GL integrand
  double UC[3][PPT], W3[PPT];    // adjoint: U . c_b and w3
#pragma unroll
  for (int q = 0; q < PPT; ++q) {
    const int p = threadIdx.x + q * NT;
    if (p >= NPG) {
#pragma unroll
      for (int c = 0; c < 3; ++c) O[c][q] = UC[c][q] = 0.0;
      W3[q] = 0.0;
      continue;
    }
    const int i = p % LXD, j = (p / LXD) % LXD, k = p / (LXD * LXD);
    const double w3 = sh.w[i] * sh.w[j] * sh.w[k];
    double c[9];
#pragma unroll
    for (int m = 0; m < 9; ++m)
      c[m] = __builtin_nontemporal_load(cof + m * ngl + gb + p);
    // opgrad of the GL field g at this point
    auto opg = [&](const double* g, double o[3]) {
      double r, s, t;
      gl_grad<LXD>(sh.Dg, g, i, j, k, r, s, t);
#pragma unroll
      for (int a = 0; a < 3; ++a)
        o[a] = w3 * (c[a] * r + c[3 + a] * s + c[6 + a] * t);
    };
    if constexpr (MODE == kFwd) {
      const double tv[3] = {sh.buf[0][p], sh.buf[1][p], sh.buf[2][p]};
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) {
        double o[3];
        opg(sh.buf[cc], o);
        O[cc][q] = o[0] * tv[0] + o[1] * tv[1] + o[2] * tv[2];
      }
    } else if constexpr (MODE == kLin) {
      const double ub[3] = {sh.buf[0][p], sh.buf[1][p], sh.buf[2][p]};
      const double up[3] = {sh.buf[3][p], sh.buf[4][p], sh.buf[5][p]};
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) {
        double o[3], o2[3];
        opg(sh.buf[cc], o);
        opg(sh.buf[3 + cc], o2);
        O[cc][q] = (o[0] * up[0] + o[1] * up[1] + o[2] * up[2]) +
                   (o2[0] * ub[0] + o2[1] * ub[1] + o2[2] * ub[2]);
      }
    } else {
      const double ub[3] = {sh.buf[0][p], sh.buf[1][p], sh.buf[2][p]};
      double o[3][3];  // o[d][c] = opgrad_c(U_d)
#pragma unroll
      for (int d = 0; d < 3; ++d) opg(sh.buf[d], o[d]);
#pragma unroll
      for (int cc = 0; cc < 3; ++cc)
        O[cc][q] = A[0][q] * o[0][cc] + A[1][q] * o[1][cc] + A[2][q] * o[2][cc];
#pragma unroll
      for (int b = 0; b < 3; ++b)
        UC[b][q] = ub[0] * c[3 * b] + ub[1] * c[3 * b + 1] + ub[2] * c[3 * b + 2];
      W3[q] = w3;
    }
  }
  __syncthreads();  // the GL input fields are consumed

#pragma unroll
  for (int cc = 0; cc < 3; ++cc) {
    if constexpr (MODE == kAdj) {
      // weak term: F_b = w3 a_c (U . c_b), then G = O_c + sum_b D_b^T F_b
#pragma unroll
      for (int q = 0; q < PPT; ++q) {
        const int p = threadIdx.x + q * NT;
        if (p < NPG) {
          const double wa = W3[q] * A[cc][q];
          sh.buf[0][p] = wa * UC[0][q];
          sh.buf[1][p] = wa * UC[1][q];
          sh.buf[2][p] = wa * UC[2][q];
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < PPT; ++q) {
        const int p = threadIdx.x + q * NT;
        if (p < NPG) {
          const int i = p % LXD, j = (p / LXD) % LXD, k = p / (LXD * LXD);
          constexpr int NS = LXD * LXD;
          double s = 0.0;
#pragma unroll
          for (int m = 0; m < LXD; ++m) {
            s += sh.Dg[m * LXD + i] * sh.buf[0][k * NS + j * LXD + m];
            s += sh.Dg[m * LXD + j] * sh.buf[1][k * NS + m * LXD + i];
            s += sh.Dg[m * LXD + k] * sh.buf[2][m * NS + j * LXD + i];
          }
          sh.buf[3][p] = O[cc][q] + s;
        }
      }
      __syncthreads();
      down_store(sh, sh.buf[3], f[cc], eb, beta);
    } else {
#pragma unroll
      for (int q = 0; q < PPT; ++q) {
        const int p = threadIdx.x + q * NT;
        if (p < NPG) sh.buf[0][p] = O[cc][q];
      }
      __syncthreads();
      down_store(sh, sh.buf[0], f[cc], eb, beta);
    }
  }
}

// GL metric: the GLL cofactors of every element (from the coordinates and
// the GLL derivative matrix) interpolated to the GL points.
template <int LX, int LXD>
__global__ void __launch_bounds__(dl_threads<LXD>())
    dealias_metric_k(double* __restrict__ cof, const double* __restrict__ x,
                     const double* __restrict__ y,
                     const double* __restrict__ z,
                     const double* __restrict__ dx,
                     const double* __restrict__ mats, int64_t ngl) {
  constexpr int NT = dl_threads<LXD>();
  constexpr int NP = LX * LX * LX, NPG = LXD * LXD * LXD, NS = LX * LX;
  __shared__ DlShared<LX, LXD, 1> sh;
  __shared__ double sx[3][NP];
  __shared__ double sc[9][NP];
  __shared__ double sdx[LX * LX];
  const int e = blockIdx.x;
  const int64_t eb = (int64_t)e * NP, gb = (int64_t)e * NPG;
  load_mats(sh, mats);
  for (int p = threadIdx.x; p < NP; p += NT) {
    sx[0][p] = x[eb + p];
    sx[1][p] = y[eb + p];
    sx[2][p] = z[eb + p];
  }
  for (int p = threadIdx.x; p < LX * LX; p += NT) sdx[p] = dx[p];
  __syncthreads();
  for (int p = threadIdx.x; p < NP; p += NT) {
    const int i = p % LX, j = (p / LX) % LX, k = p / NS;
    double dr[3] = {0, 0, 0}, ds[3] = {0, 0, 0}, dt[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int m = 0; m < LX; ++m) {
        dr[a] += sdx[i * LX + m] * sx[a][k * NS + j * LX + m];
        ds[a] += sdx[j * LX + m] * sx[a][k * NS + m * LX + i];
        dt[a] += sdx[k * LX + m] * sx[a][m * NS + j * LX + i];
      }
    // c_r = x_s x x_t, c_s = x_t x x_r, c_t = x_r x x_s
    sc[0][p] = ds[1] * dt[2] - ds[2] * dt[1];
    sc[1][p] = ds[2] * dt[0] - ds[0] * dt[2];
    sc[2][p] = ds[0] * dt[1] - ds[1] * dt[0];
    sc[3][p] = dt[1] * dr[2] - dt[2] * dr[1];
    sc[4][p] = dt[2] * dr[0] - dt[0] * dr[2];
    sc[5][p] = dt[0] * dr[1] - dt[1] * dr[0];
    sc[6][p] = dr[1] * ds[2] - dr[2] * ds[1];
    sc[7][p] = dr[2] * ds[0] - dr[0] * ds[2];
    sc[8][p] = dr[0] * ds[1] - dr[1] * ds[0];
  }
  __syncthreads();
  for (int m = 0; m < 9; ++m) {
    for (int p = threadIdx.x; p < NP; p += NT) sh.t2[p] = sc[m][p];
    __syncthreads();
    interp_up(sh, sh.buf[0]);
    for (int p = threadIdx.x; p < NPG; p += NT)
      cof[m * ngl + gb + p] = sh.buf[0][p];
    __syncthreads();
  }
}

// ---- host: the 1-D rules -------------------------------------------------

// Gauss-Legendre nodes (ascending) and weights: Newton on the three-term
// recurrence of P_n.
void gauss_legendre(int n, std::vector<double>& z, std::vector<double>& w) {
  z.assign(n, 0.0);
  w.assign(n, 0.0);
  auto legendre = [n](double x, double& pn, double& dpn) {
    double p0 = 1.0, p1 = x;
    for (int k = 2; k <= n; ++k) {
      const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    pn = p1;
    dpn = n * (x * p1 - p0) / (x * x - 1.0);
  };
  for (int i = 0; i < n; ++i) {
    double x = std::cos(M_PI * (i + 0.75) / (n + 0.5)), pn, dpn;
    for (int it = 0; it < 100; ++it) {
      legendre(x, pn, dpn);
      const double step = pn / dpn;
      x -= step;
      if (std::fabs(step) < 1e-16) break;
    }
    legendre(x, pn, dpn);
    z[n - 1 - i] = x;
    w[n - 1 - i] = 2.0 / ((1.0 - x * x) * dpn * dpn);
  }
}

template <int LX, int LXD>
int dealias_run(const DealiasState* s, int mode, double* fx, double* fy,
                double* fz, const double* a0, const double* a1,
                const double* a2, const double* b0, const double* b1,
                const double* b2, double beta) {
  const dim3 grid(s->nelv), block(dl_threads<LXD>());
  switch (mode) {
    case kFwd:
      hipLaunchKernelGGL((dealias_k<LX, LXD, kFwd>), grid, block, 0, stream(),
                         fx, fy, fz, a0, a1, a2, b0, b1, b2, s->cof, s->mats,
                         s->ngl, beta);
      break;
    case kLin:
      hipLaunchKernelGGL((dealias_k<LX, LXD, kLin>), grid, block, 0, stream(),
                         fx, fy, fz, a0, a1, a2, b0, b1, b2, s->cof, s->mats,
                         s->ngl, beta);
      break;
    default:
      hipLaunchKernelGGL((dealias_k<LX, LXD, kAdj>), grid, block, 0, stream(),
                         fx, fy, fz, a0, a1, a2, b0, b1, b2, s->cof, s->mats,
                         s->ngl, beta);
  }
  NK_LAUNCHED();
  return kOk;
}

#define NK_DL_SWITCH(LXV, CALL_T)                  \
  switch (LXV) {                                   \
    case 4: { constexpr int L = 4, LD = 6; CALL_T; } \
    case 6: { constexpr int L = 6, LD = 9; CALL_T; } \
    case 8: { constexpr int L = 8, LD = 12; CALL_T; } \
    default: break;                                \
  }

int dealias_apply(void* st, int mode, void* fx, void* fy, void* fz,
                  const void* a0, const void* a1, const void* a2,
                  const void* b0, const void* b1, const void* b2,
                  double beta) {
  NK_INIT();
  auto* s = static_cast<DealiasState*>(st);
  NK_REQUIRE(s != nullptr, "null dealias state");
  if (s->nelv == 0) return kOk;
  NK_REQUIRE(fx && fy && fz && a0 && a1 && a2, "null field");
  NK_REQUIRE(mode == kFwd || (b0 && b1 && b2), "null base flow");
  auto* X = (double*)fx;
  auto* Y = (double*)fy;
  auto* Z = (double*)fz;
  auto* A0 = (const double*)a0;
  auto* A1 = (const double*)a1;
  auto* A2 = (const double*)a2;
  auto* B0 = (const double*)b0;
  auto* B1 = (const double*)b1;
  auto* B2 = (const double*)b2;
  NK_DL_SWITCH(s->lx, return (dealias_run<L, LD>(s, mode, X, Y, Z, A0, A1, A2,
                                                  B0, B1, B2, beta)));
  return fail_arg(__func__, "unsupported lx");
}

}  // namespace
}  // namespace nekoamd

using namespace nekoamd;

extern "C" {

int nekoamd_dealias_create(void** st, const void* x, const void* y,
                           const void* z, const void* dx, const int* lx,
                           const int* lxd, const int* nelv) {
  NK_INIT();
  NK_REQUIRE(st && lx && lxd && nelv, "null argument");
  NK_REQUIRE(*nelv >= 0, "nelv < 0");
  if (!(*lx == 4 || *lx == 6 || *lx == 8) || *lxd != 3 * *lx / 2) {
    set_error(__func__, "dealiasing supports lx in {4, 6, 8} with lxd = 3 lx / 2");
    return kErrUnsupported;
  }
  NK_REQUIRE(*nelv == 0 || (x && y && z && dx), "null coordinates");
  auto* s = new DealiasState;
  s->lx = *lx;
  s->lxd = *lxd;
  s->nelv = *nelv;
  s->n = (int64_t)*nelv * *lx * *lx * *lx;
  s->ngl = (int64_t)*nelv * *lxd * *lxd * *lxd;
  const int L = *lx, LD = *lxd;
  std::vector<double> zl(L), wl(L), dl(L * L), zg, wg;
  int rc = nekoamd_gll(lx, zl.data(), wl.data(), dl.data());
  if (rc != kOk) {
    delete s;
    return rc;
  }
  gauss_legendre(LD, zg, wg);
  // J[q*L + i] = l_i(zg_q); Dg[i*LD + j] = l'_j(zg_i) on the GL nodes
  std::vector<double> mats(LD * L + LD * LD + LD);
  for (int q = 0; q < LD; ++q)
    for (int i = 0; i < L; ++i) {
      double v = 1.0;
      for (int k = 0; k < L; ++k)
        if (k != i) v *= (zg[q] - zl[k]) / (zl[i] - zl[k]);
      mats[q * L + i] = v;
    }
  std::vector<double> a(LD, 1.0);
  for (int i = 0; i < LD; ++i)
    for (int k = 0; k < LD; ++k)
      if (k != i) a[i] *= zg[i] - zg[k];
  double* Dg = mats.data() + LD * L;
  for (int i = 0; i < LD; ++i) {
    double d = 0.0;
    for (int j = 0; j < LD; ++j) {
      if (j == i) continue;
      Dg[i * LD + j] = (a[i] / a[j]) / (zg[i] - zg[j]);
      d += 1.0 / (zg[i] - zg[j]);
    }
    Dg[i * LD + i] = d;
  }
  for (int i = 0; i < LD; ++i) mats[LD * L + LD * LD + i] = wg[i];
  hipError_t e = hipMalloc(&s->mats, sizeof(double) * mats.size());
  if (e == hipSuccess && s->ngl > 0)
    e = hipMalloc(&s->cof, sizeof(double) * 9 * (size_t)s->ngl);
  if (e == hipSuccess)
    e = hipMemcpyAsync(s->mats, mats.data(), sizeof(double) * mats.size(),
                       hipMemcpyHostToDevice, stream());
  if (e == hipSuccess && s->nelv > 0) {
    auto* X = (const double*)x;
    auto* Y = (const double*)y;
    auto* Z = (const double*)z;
    auto* D = (const double*)dx;
    NK_DL_SWITCH(L, hipLaunchKernelGGL((dealias_metric_k<L, LD>),
                                       dim3(s->nelv), dim3(dl_threads<LD>()),
                                       0, stream(), s->cof, X, Y, Z, D,
                                       s->mats, s->ngl);
                 break);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(stream());
  if (e != hipSuccess) {
    if (s->cof) (void)hipFree(s->cof);
    if (s->mats) (void)hipFree(s->mats);
    delete s;
    return fail_hip(__func__, e);
  }
  *st = s;
  return kOk;
}

int nekoamd_dealias_free(void* st) {
  NK_INIT();
  auto* s = static_cast<DealiasState*>(st);
  if (!s) return kOk;
  NK_HIP(hipStreamSynchronize(stream()));
  if (s->cof) NK_HIP(hipFree(s->cof));
  if (s->mats) NK_HIP(hipFree(s->mats));
  delete s;
  return kOk;
}

int nekoamd_dealias_metric(const void* st, void* cof) {
  NK_INIT();
  auto* s = static_cast<const DealiasState*>(st);
  NK_REQUIRE(s && (cof || s->ngl == 0), "null argument");
  if (s->ngl == 0) return kOk;
  NK_HIP(hipMemcpyAsync(cof, s->cof, sizeof(double) * 9 * (size_t)s->ngl,
                        hipMemcpyDeviceToDevice, stream()));
  return kOk;
}

int nekoamd_dealias_advection(void* st, void* fx, void* fy, void* fz,
                              const void* u, const void* v, const void* w,
                              const double* beta) {
  NK_REQUIRE(beta != nullptr, "null beta");
  return dealias_apply(st, kFwd, fx, fy, fz, u, v, w, nullptr, nullptr,
                       nullptr, *beta);
}

int nekoamd_dealias_linear(void* st, void* fx, void* fy, void* fz,
                           const void* up, const void* vp, const void* wp,
                           const void* ub, const void* vb, const void* wb) {
  return dealias_apply(st, kLin, fx, fy, fz, up, vp, wp, ub, vb, wb, 1.0);
}

int nekoamd_dealias_adjoint(void* st, void* fx, void* fy, void* fz,
                            const void* ua, const void* va, const void* wa,
                            const void* ub, const void* vb, const void* wb) {
  return dealias_apply(st, kAdj, fx, fy, fz, ua, va, wa, ub, vb, wb, 1.0);
}

}  // extern "C"
