// pnpn.hip -- the operators of one Pn/Pn time step around the solves
// (sources/adjoint/adjoint_pnpn.f90:619-827 calls them through Neko's
// sumab, makeabf, makebdf, adv, pnpn_prs_res, pnpn_vel_res):
//
//   advection     B f = -B (u.grad) u                  element pass, fused x3
//   curl          element-local curl (averaged by the caller with gs)
//   opgrad        B grad p = w3 sum_b c_b p_b           element pass
//   cdtp          sum_a sum_b D_b^T (w3 c_b[a] t_a)     element pass (weak div)
//   surface flux  int_face q u.n on flagged element faces
//   fused elementwise stages: EXT of u, makeabf + makebdf + lag update,
//   pressure forcing, pressure and velocity residual combinations.
//
// The derivative kernels (advection, curl, opgrad, cdtp) are one workgroup
// per element and one thread per point, coordinates, fields and D staged in
// LDS (the surface flux: one thread per (i,j) column walking k); the metric
// cofactors are recomputed from the coordinates (3 fields read instead of 9
// stored).  Measured at 32^3 p=7 against the column form: curl 295 against
// 314 us, advection 321 against 384, cdtp 246 against 386, opgrad 199
// against 173.
#include "internal.h"
#include "nekoamd.h"

namespace nekoamd {

constexpr int kPB = 256;

template <typename F>
__global__ void __launch_bounds__(kPB) pmap_k(int64_t n, F f) {
  for (int64_t i = (int64_t)blockIdx.x * kPB + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kPB)
    f(i);
}

template <typename F>
static int pmap(int64_t n, F f) {
  if (n <= 0) return kOk;
  const int g = nk_grid(n, kPB, kMapGridCap);
  hipLaunchKernelGGL((pmap_k<F>), dim3(g), dim3(kPB), 0,
                     stream(), n, f);
  NK_LAUNCHED();
  return kOk;
}

// r/s/t derivative of LDS field a at (i,j,k).
template <int LX>
__device__ __forceinline__ void deriv3(const double* __restrict__ dx,
                                       const double* a, int i, int j, int k,
                                       double& r, double& s, double& t) {
  constexpr int NS = LX * LX;
  r = 0.0;
  s = 0.0;
  t = 0.0;
#pragma unroll
  for (int m = 0; m < LX; ++m) {
    r += dx[i * LX + m] * a[k * NS + j * LX + m];
    s += dx[j * LX + m] * a[k * NS + m * LX + i];
    t += dx[k * LX + m] * a[m * NS + j * LX + i];
  }
}

// deriv3 for the one-thread-per-point kernels: D rows i and j from LDS (sD),
// row k from global memory (dx).  At lx = 8 a wavefront is one k-plane, so
// k is made wave-uniform and row k arrives by scalar loads, taking a third
// of the derivative's LDS reads off the LDS.  Same sums in the same order
// as deriv3.
template <int LX>
__device__ __forceinline__ void deriv3p(const double* sD,
                                        const double* __restrict__ dx,
                                        const double* a, int i, int j, int k,
                                        double& r, double& s, double& t) {
  constexpr int NS = LX * LX;
  const int kr = NS == 64 ? __builtin_amdgcn_readfirstlane(k) : k;
  const double* Dk = NS == 64 ? dx : sD;
  r = 0.0;
  s = 0.0;
  t = 0.0;
#pragma unroll
  for (int m = 0; m < LX; ++m) {
    r += sD[i * LX + m] * a[k * NS + j * LX + m];
    s += sD[j * LX + m] * a[k * NS + m * LX + i];
    t += Dk[kr * LX + m] * a[m * NS + j * LX + i];
  }
}

struct Cof9 {
  double c[9];  // c_r (0..2), c_s (3..5), c_t (6..8)
  double J;
};

template <int LX>
__device__ __forceinline__ Cof9 cof_at(const double* __restrict__ dx,
                                       const double* sx, const double* sy,
                                       const double* sz, int i, int j, int k) {
  double xr, xs, xt, yr, ys, yt, zr, zs, zt;
  deriv3<LX>(dx, sx, i, j, k, xr, xs, xt);
  deriv3<LX>(dx, sy, i, j, k, yr, ys, yt);
  deriv3<LX>(dx, sz, i, j, k, zr, zs, zt);
  Cof9 o;
  o.c[0] = ys * zt - zs * yt;
  o.c[1] = zs * xt - xs * zt;
  o.c[2] = xs * yt - ys * xt;
  o.c[3] = yt * zr - zt * yr;
  o.c[4] = zt * xr - xt * zr;
  o.c[5] = xt * yr - yt * xr;
  o.c[6] = yr * zs - zr * ys;
  o.c[7] = zr * xs - xr * zs;
  o.c[8] = xr * ys - yr * xs;
  o.J = xr * o.c[0] + yr * o.c[1] + zr * o.c[2];
  return o;
}

template <int LX>
__device__ __forceinline__ Cof9 cof_at_p(const double* sD,
                                         const double* __restrict__ dx,
                                         const double* sx, const double* sy,
                                         const double* sz, int i, int j,
                                         int k) {
  double xr, xs, xt, yr, ys, yt, zr, zs, zt;
  deriv3p<LX>(sD, dx, sx, i, j, k, xr, xs, xt);
  deriv3p<LX>(sD, dx, sy, i, j, k, yr, ys, yt);
  deriv3p<LX>(sD, dx, sz, i, j, k, zr, zs, zt);
  Cof9 o;
  o.c[0] = ys * zt - zs * yt;
  o.c[1] = zs * xt - xs * zt;
  o.c[2] = xs * yt - ys * xt;
  o.c[3] = yt * zr - zt * yr;
  o.c[4] = zt * xr - xt * zr;
  o.c[5] = xt * yr - yt * xr;
  o.c[6] = yr * zs - zr * ys;
  o.c[7] = zr * xs - xr * zs;
  o.c[8] = xr * ys - yr * xs;
  o.J = xr * o.c[0] + yr * o.c[1] + zr * o.c[2];
  return o;
}

// MODE 0: advection  out_a = -B (u.grad) u_a          (in: u,v,w; B)
// MODE 1: curl       out = curl(u,v,w)                 (in: u,v,w)
// MODE 2: opgrad     out_a = w3 sum_b c_b[a] p_b       (in: p)
// One thread per point (t = i + LX (j + LX k)), the element's coordinates,
// fields and D in LDS.  (One thread per (i, j) column walking k held 64
// threads and 24 KB of LDS per element: 1.5 wavefronts per SIMD.)
// Advection and curl at lx = 8 are held to 64 VGPRs (8 waves per SIMD, 4
// elements per CU instead of 2; no spill, the curl summed field by field):
// advection 302 against 323 us, curl 291 against 296 us at 32^3 p=7; opgrad
// already fits.
template <int LX, int MODE>
__global__ void __launch_bounds__(LX * LX * LX, LX == 8 && MODE != 2 ? 8 : 1)
    deriv_k(double* __restrict__ o0, double* __restrict__ o1,
            double* __restrict__ o2, const double* __restrict__ f0,
            const double* __restrict__ f1, const double* __restrict__ f2,
            const double* __restrict__ xc, const double* __restrict__ yc,
            const double* __restrict__ zc, const double* __restrict__ dx,
            const double* __restrict__ wq, const double* __restrict__ bm) {
  constexpr int NS = LX * LX, NP = NS * LX;
  constexpr int NF = MODE == 2 ? 1 : 3;
  __shared__ double s[3 + NF][NP];
  __shared__ double sD[NS];
  const int e = blockIdx.x, t = threadIdx.x;
  const int i = t % LX, j = (t / LX) % LX, k = t / NS;
  const int64_t g = (int64_t)e * NP + t;
  if (t < NS) sD[t] = dx[t];
  const double* src[6] = {xc, yc, zc, f0, f1, f2};
#pragma unroll
  for (int a = 0; a < 3 + NF; ++a) s[a][t] = src[a][g];
  double b = 0.0;
  if constexpr (MODE == 0) b = bm[g];
  __syncthreads();
  const Cof9 c = cof_at_p<LX>(sD, dx, s[0], s[1], s[2], i, j, k);
  if constexpr (MODE == 2) {
    double pr, ps, pt;
    deriv3p<LX>(sD, dx, s[3], i, j, k, pr, ps, pt);
    const double w3 = wq[i] * wq[j] * wq[k];
    o0[g] = w3 * (c.c[0] * pr + c.c[3] * ps + c.c[6] * pt);
    o1[g] = w3 * (c.c[1] * pr + c.c[4] * ps + c.c[7] * pt);
    o2[g] = w3 * (c.c[2] * pr + c.c[5] * ps + c.c[8] * pt);
  } else {
    if constexpr (MODE == 1) {
      // curl accumulated field by field (o0 = G21 - G12, o1 = G02 - G20,
      // o2 = G10 - G01 with G[a][b] = d f_a / d x_b; the same values bit for
      // bit), so only the three sums stay live across the fields
      double cx = 0.0, cy = 0.0, cz = 0.0;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        double fr, fs, ft;
        deriv3p<LX>(sD, dx, s[3 + a], i, j, k, fr, fs, ft);
        const int q1 = (a + 1) % 3, q2 = (a + 2) % 3;
        const double g1 = (c.c[q1] * fr + c.c[3 + q1] * fs + c.c[6 + q1] * ft) / c.J;
        const double g2 = (c.c[q2] * fr + c.c[3 + q2] * fs + c.c[6 + q2] * ft) / c.J;
        // f_a contributes +G[a][a+1] to component a+2, -G[a][a+2] to a+1
        if (a == 0) { cz = -g1; cy = g2; }
        if (a == 1) { cx = -g1; cz = cz + g2; }
        if (a == 2) { cy = cy - g1; cx = cx + g2; }
      }
      o0[g] = cx;
      o1[g] = cy;
      o2[g] = cz;
      return;
    }
    double G[3][3];  // G[a][b] = d f_a / d x_b
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      double fr, fs, ft;
      deriv3p<LX>(sD, dx, s[3 + a], i, j, k, fr, fs, ft);
#pragma unroll
      for (int q = 0; q < 3; ++q)
        G[a][q] = (c.c[q] * fr + c.c[3 + q] * fs + c.c[6 + q] * ft) / c.J;
    }
    if constexpr (MODE == 0) {
      const double u = s[3][t], v = s[4][t], w = s[5][t];
      o0[g] = -b * (u * G[0][0] + v * G[0][1] + w * G[0][2]);
      o1[g] = -b * (u * G[1][0] + v * G[1][1] + w * G[1][2]);
      o2[g] = -b * (u * G[2][0] + v * G[2][1] + w * G[2][2]);
    } else {
      o0[g] = G[2][1] - G[1][2];
      o1[g] = G[0][2] - G[2][0];
      o2[g] = G[1][0] - G[0][1];
    }
  }
}

// Weak divergence: out = sum_b D_b^T g_b, g_b = w3 sum_a c_b[a] t_a.  One
// thread per point, as deriv_k.
// With binv: t_a scaled by binv first (Neko's col2(t_a, Binv) before
// cdtp); with sb: out = cdtp - sb - S sc (sub2s after it).
template <int LX>
__global__ void __launch_bounds__(LX * LX * LX)
    cdtp_k(double* __restrict__ out, const double* __restrict__ tx,
           const double* __restrict__ ty, const double* __restrict__ tz,
           const double* __restrict__ xc, const double* __restrict__ yc,
           const double* __restrict__ zc, const double* __restrict__ dx,
           const double* __restrict__ wq, const double* __restrict__ binv,
           const double* __restrict__ sb, const double* __restrict__ sc,
           double S) {
  constexpr int NS = LX * LX, NP = NS * LX;
  __shared__ double sx[3][NP];
  __shared__ double sg[3][NP];
  __shared__ double sD[NS];
  const int e = blockIdx.x, t = threadIdx.x;
  const int i = t % LX, j = (t / LX) % LX, k = t / NS;
  const int64_t g = (int64_t)e * NP + t;
  if (t < NS) sD[t] = dx[t];
  sx[0][t] = xc[g];
  sx[1][t] = yc[g];
  sx[2][t] = zc[g];
  double a0 = tx[g], a1 = ty[g], a2 = tz[g];
  if (binv != nullptr) {
    const double bi = binv[g];
    a0 = a0 * bi;
    a1 = a1 * bi;
    a2 = a2 * bi;
  }
  __syncthreads();
  const Cof9 c = cof_at_p<LX>(sD, dx, sx[0], sx[1], sx[2], i, j, k);
  const double w3 = wq[i] * wq[j] * wq[k];
  sg[0][t] = w3 * (c.c[0] * a0 + c.c[1] * a1 + c.c[2] * a2);
  sg[1][t] = w3 * (c.c[3] * a0 + c.c[4] * a1 + c.c[5] * a2);
  sg[2][t] = w3 * (c.c[6] * a0 + c.c[7] * a1 + c.c[8] * a2);
  __syncthreads();
  // column k of D: by scalar loads at lx = 8 (k is wave-uniform), as deriv3p
  const int kr = NS == 64 ? __builtin_amdgcn_readfirstlane(k) : k;
  const double* Dk = NS == 64 ? dx : sD;
  double acc = 0.0;
#pragma unroll
  for (int m = 0; m < LX; ++m) {
    acc += sD[m * LX + i] * sg[0][k * NS + j * LX + m];
    acc += sD[m * LX + j] * sg[1][k * NS + m * LX + i];
    acc += Dk[m * LX + kr] * sg[2][m * NS + j * LX + i];
  }
  if (sb != nullptr)
    out[g] = acc - sb[g] - S * sc[g];
  else
    out[g] = acc;
}

// Boundary flux on the faces flagged in fbits[e] (bit f: face f = 2*dir +
// hi, dir 0/1/2 = r/s/t): out(pt) = sum_faces w w (+-c_dir . u).  out is
// written for every point of the element (zero off the flagged faces).
template <int LX>
__global__ void __launch_bounds__(LX * LX)
    surf_k(double* __restrict__ out, const double* __restrict__ u,
           const double* __restrict__ v, const double* __restrict__ w,
           const double* __restrict__ xc, const double* __restrict__ yc,
           const double* __restrict__ zc, const double* __restrict__ dx,
           const double* __restrict__ wq, const int* __restrict__ fbits) {
  constexpr int NS = LX * LX, NP = NS * LX, P = LX - 1;
  __shared__ double sx[3][NP];
  __shared__ double acc[NP];
  const int e = blockIdx.x, t = threadIdx.x;
  const int64_t eb = (int64_t)e * NP;
  const int bits = fbits[e];
#pragma unroll
  for (int k = 0; k < LX; ++k) {
    acc[k * NS + t] = 0.0;
    if (bits) {
      sx[0][k * NS + t] = xc[eb + k * NS + t];
      sx[1][k * NS + t] = yc[eb + k * NS + t];
      sx[2][k * NS + t] = zc[eb + k * NS + t];
    }
  }
  __syncthreads();
  if (bits) {
    // faces one after the other so shared edge points accumulate in order
    for (int f = 0; f < 6; ++f) {
      if (!((bits >> f) & 1)) continue;
      const int dir = f >> 1, hi = f & 1;
      const int a = t % LX, b = t / LX;
      int i, j, k;
      double wt;
      if (dir == 0) { i = hi ? P : 0; j = a; k = b; wt = wq[j] * wq[k]; }
      else if (dir == 1) { j = hi ? P : 0; i = a; k = b; wt = wq[i] * wq[k]; }
      else { k = hi ? P : 0; i = a; j = b; wt = wq[i] * wq[j]; }
      const Cof9 c = cof_at<LX>(dx, sx[0], sx[1], sx[2], i, j, k);
      const int q = i + LX * (j + LX * k);
      const double un = c.c[3 * dir] * u[eb + q] + c.c[3 * dir + 1] * v[eb + q] +
                        c.c[3 * dir + 2] * w[eb + q];
      acc[q] += (hi ? 1.0 : -1.0) * wt * un;
      __syncthreads();
    }
  }
#pragma unroll
  for (int k = 0; k < LX; ++k) out[eb + k * NS + t] = acc[k * NS + t];
}

#define NK_LX_SWITCH(LXV, CALL)                                  \
  switch (LXV) {                                                 \
    case 2: { constexpr int L = 2; CALL; } break;               \
    case 3: { constexpr int L = 3; CALL; } break;               \
    case 4: { constexpr int L = 4; CALL; } break;               \
    case 5: { constexpr int L = 5; CALL; } break;               \
    case 6: { constexpr int L = 6; CALL; } break;               \
    case 7: { constexpr int L = 7; CALL; } break;               \
    case 8: { constexpr int L = 8; CALL; } break;               \
    default: return fail_arg(__func__, "lx must be in [2, 8]"); \
  }

}  // namespace nekoamd

using namespace nekoamd;

#define NK_PN(n) \
  NK_INIT();     \
  NK_REQUIRE(n && *n >= 0, "bad length")

extern "C" {

int nekoamd_advection(void* fx, void* fy, void* fz, const void* u,
                      const void* v, const void* w, const void* B,
                      const void* x, const void* y, const void* z,
                      const void* dx, const int* lx, const int* nelv) {
  NK_INIT();
  NK_REQUIRE(fx && fy && fz && u && v && w && B && x && y && z && dx && lx &&
                 nelv,
             "null argument");
  if (*nelv <= 0) return kOk;
  NK_LX_SWITCH(*lx, hipLaunchKernelGGL((deriv_k<L, 0>), dim3(*nelv),
                                       dim3(L * L * L), 0, stream(), (double*)fx,
                                       (double*)fy, (double*)fz,
                                       (const double*)u, (const double*)v,
                                       (const double*)w, (const double*)x,
                                       (const double*)y, (const double*)z,
                                       (const double*)dx, nullptr,
                                       (const double*)B));
  NK_LAUNCHED();
  return kOk;
}

int nekoamd_curl(void* cx, void* cy, void* cz, const void* u, const void* v,
                 const void* w, const void* x, const void* y, const void* z,
                 const void* dx, const int* lx, const int* nelv) {
  NK_INIT();
  NK_REQUIRE(cx && cy && cz && u && v && w && x && y && z && dx && lx && nelv,
             "null argument");
  if (*nelv <= 0) return kOk;
  NK_LX_SWITCH(*lx, hipLaunchKernelGGL((deriv_k<L, 1>), dim3(*nelv),
                                       dim3(L * L * L), 0, stream(), (double*)cx,
                                       (double*)cy, (double*)cz,
                                       (const double*)u, (const double*)v,
                                       (const double*)w, (const double*)x,
                                       (const double*)y, (const double*)z,
                                       (const double*)dx, nullptr, nullptr));
  NK_LAUNCHED();
  return kOk;
}

int nekoamd_opgrad(void* ox, void* oy, void* oz, const void* p, const void* x,
                   const void* y, const void* z, const void* dx,
                   const void* wgll, const int* lx, const int* nelv) {
  NK_INIT();
  NK_REQUIRE(ox && oy && oz && p && x && y && z && dx && wgll && lx && nelv,
             "null argument");
  if (*nelv <= 0) return kOk;
  NK_LX_SWITCH(*lx, hipLaunchKernelGGL((deriv_k<L, 2>), dim3(*nelv),
                                       dim3(L * L * L), 0, stream(), (double*)ox,
                                       (double*)oy, (double*)oz,
                                       (const double*)p, nullptr, nullptr,
                                       (const double*)x, (const double*)y,
                                       (const double*)z, (const double*)dx,
                                       (const double*)wgll, nullptr));
  NK_LAUNCHED();
  return kOk;
}

int nekoamd_cdtp(void* out, const void* tx, const void* ty, const void* tz,
                 const void* x, const void* y, const void* z, const void* dx,
                 const void* wgll, const int* lx, const int* nelv) {
  NK_INIT();
  NK_REQUIRE(out && tx && ty && tz && x && y && z && dx && wgll && lx && nelv,
             "null argument");
  if (*nelv <= 0) return kOk;
  NK_LX_SWITCH(*lx, hipLaunchKernelGGL((cdtp_k<L>), dim3(*nelv), dim3(L * L * L),
                                       0, stream(), (double*)out,
                                       (const double*)tx, (const double*)ty,
                                       (const double*)tz, (const double*)x,
                                       (const double*)y, (const double*)z,
                                       (const double*)dx, (const double*)wgll,
                                       nullptr, nullptr, nullptr, 0.0));
  NK_LAUNCHED();
  return kOk;
}

int nekoamd_pnpn_prs_div(void* out, const void* tx, const void* ty,
                         const void* tz, const void* binv, const void* b,
                         const void* c, const double* s, const void* x,
                         const void* y, const void* z, const void* dx,
                         const void* wgll, const int* lx, const int* nelv) {
  NK_INIT();
  NK_REQUIRE(out && tx && ty && tz && binv && b && c && s && x && y && z &&
                 dx && wgll && lx && nelv,
             "null argument");
  if (*nelv <= 0) return kOk;
  NK_LX_SWITCH(*lx, hipLaunchKernelGGL((cdtp_k<L>), dim3(*nelv), dim3(L * L * L),
                                       0, stream(), (double*)out,
                                       (const double*)tx, (const double*)ty,
                                       (const double*)tz, (const double*)x,
                                       (const double*)y, (const double*)z,
                                       (const double*)dx, (const double*)wgll,
                                       (const double*)binv, (const double*)b,
                                       (const double*)c, *s));
  NK_LAUNCHED();
  return kOk;
}

int nekoamd_adjoint_forcing(void* fx, void* fy, void* fz, const void* cx,
                            const void* cy, const void* cz, const void* ub,
                            const void* vb, const void* wb, const void* chi,
                            const void* B, const double* K,
                            const double* obj_scale, const int64_t* n) {
  NK_INIT();
  NK_REQUIRE(fx && fy && fz && cx && cy && cz && ub && vb && wb && chi && B &&
                 K && obj_scale && n,
             "null argument");
  double* Fx = (double*)fx;
  double* Fy = (double*)fy;
  double* Fz = (double*)fz;
  const double *Cx = (const double*)cx, *Cy = (const double*)cy,
               *Cz = (const double*)cz, *Ub = (const double*)ub,
               *Vb = (const double*)vb, *Wb = (const double*)wb,
               *Ch = (const double*)chi, *Bm = (const double*)B;
  const double k = *K, os = *obj_scale;
  auto f = [=] __device__(int64_t i) {
    const double kc = k * Ch[i];
    Fx[i] = (os * Cx[i] + kc * Ub[i]) * Bm[i];
    Fy[i] = (os * Cy[i] + kc * Vb[i]) * Bm[i];
    Fz[i] = (os * Cz[i] + kc * Wb[i]) * Bm[i];
  };
  return pmap(*n, f);
}

int nekoamd_adjoint_advection(void* fx, void* fy, void* fz, const void* ua,
                              const void* va, const void* wa, const void* ub,
                              const void* vb, const void* wb, const void* x,
                              const void* y, const void* z, const void* dx,
                              const void* wgll, const int* lx,
                              const int* nelv, void* work) {
  NK_INIT();
  NK_REQUIRE(fx && fy && fz && ua && va && wa && ub && vb && wb && x && y &&
                 z && dx && wgll && lx && nelv && work,
             "null argument");
  NK_REQUIRE(*lx >= 2 && *lx <= NEKOAMD_LX_MAX && *nelv >= 0, "bad size");
  const int64_t n = (int64_t)*nelv * (*lx) * (*lx) * (*lx);
  if (n == 0) return kOk;
  double* g = (double*)work;
  double* pr = g + 9 * n;
  const void* base[3] = {ub, vb, wb};
  for (int d = 0; d < 3; ++d) {
    int rc = nekoamd_opgrad(g + 3 * d * n, g + (3 * d + 1) * n,
                            g + (3 * d + 2) * n, base[d], x, y, z, dx, wgll,
                            lx, nelv);
    if (rc) return rc;
  }
  double* Fx = (double*)fx;
  double* Fy = (double*)fy;
  double* Fz = (double*)fz;
  const double *Ua = (const double*)ua, *Va = (const double*)va,
               *Wa = (const double*)wa;
  const double* G = g;
  auto dotk = [=] __device__(int64_t i) {
    Fx[i] -= Ua[i] * G[i] + Va[i] * G[3 * n + i] + Wa[i] * G[6 * n + i];
    Fy[i] -= Ua[i] * G[n + i] + Va[i] * G[4 * n + i] + Wa[i] * G[7 * n + i];
    Fz[i] -= Ua[i] * G[2 * n + i] + Va[i] * G[5 * n + i] + Wa[i] * G[8 * n + i];
  };
  int rc = pmap(n, dotk);
  if (rc) return rc;
  const double* Ub = (const double*)ub;
  const double* Vb = (const double*)vb;
  const double* Wb = (const double*)wb;
  const double* uc[3] = {Ua, Va, Wa};
  double* fc[3] = {Fx, Fy, Fz};
  for (int c = 0; c < 3; ++c) {
    const double* Uc = uc[c];
    auto prod = [=] __device__(int64_t i) {
      pr[i] = Uc[i] * Ub[i];
      pr[n + i] = Uc[i] * Vb[i];
      pr[2 * n + i] = Uc[i] * Wb[i];
    };
    rc = pmap(n, prod);
    if (!rc)
      rc = nekoamd_cdtp(g, pr, pr + n, pr + 2 * n, x, y, z, dx, wgll, lx,
                        nelv);
    if (rc) return rc;
    double* Fc = fc[c];
    auto sub = [=] __device__(int64_t i) { Fc[i] -= G[i]; };
    rc = pmap(n, sub);
    if (rc) return rc;
  }
  return kOk;
}

int nekoamd_surface_flux(void* out, const void* u, const void* v,
                         const void* w, const void* x, const void* y,
                         const void* z, const void* dx, const void* wgll,
                         const void* face_bits, const int* lx,
                         const int* nelv) {
  NK_INIT();
  NK_REQUIRE(out && u && v && w && x && y && z && dx && wgll && face_bits &&
                 lx && nelv,
             "null argument");
  if (*nelv <= 0) return kOk;
  NK_LX_SWITCH(*lx, hipLaunchKernelGGL((surf_k<L>), dim3(*nelv), dim3(L * L), 0,
                                       stream(), (double*)out, (const double*)u,
                                       (const double*)v, (const double*)w,
                                       (const double*)x, (const double*)y,
                                       (const double*)z, (const double*)dx,
                                       (const double*)wgll,
                                       (const int*)face_bits));
  NK_LAUNCHED();
  return kOk;
}

// Element face bits of the box faces in `faces` (host helper).
int nekoamd_box_face_bits(const int* nx, const int* ny, const int* nz,
                          const int* faces, int* bits) {
  NK_REQUIRE(nx && ny && nz && faces && bits, "null argument");
  const int f = *faces;
  int64_t e = 0;
  for (int ez = 0; ez < *nz; ++ez)
    for (int ey = 0; ey < *ny; ++ey)
      for (int ex = 0; ex < *nx; ++ex, ++e) {
        int b = 0;
        if ((f & NEKOAMD_FACE_XMIN) && ex == 0) b |= 1;
        if ((f & NEKOAMD_FACE_XMAX) && ex == *nx - 1) b |= 2;
        if ((f & NEKOAMD_FACE_YMIN) && ey == 0) b |= 4;
        if ((f & NEKOAMD_FACE_YMAX) && ey == *ny - 1) b |= 8;
        if ((f & NEKOAMD_FACE_ZMIN) && ez == 0) b |= 16;
        if ((f & NEKOAMD_FACE_ZMAX) && ez == *nz - 1) b |= 32;
        bits[e] = b;
      }
  return kOk;
}

// ---- fused elementwise stages ----------------------------------------------

// sumab: ue = ext1 u + ext2 ulag for three components.
int nekoamd_pnpn_extrapolate(void* ue, void* ve, void* we, const void* u,
                             const void* v, const void* w, const void* ul,
                             const void* vl, const void* wl,
                             const double* ext, const int64_t* n) {
  NK_PN(n);
  NK_REQUIRE(ext, "null coefficients");
  double *Ue = (double*)ue, *Ve = (double*)ve, *We = (double*)we;
  const double *U = (const double*)u, *V = (const double*)v,
               *W = (const double*)w, *Ul = (const double*)ul,
               *Vl = (const double*)vl, *Wl = (const double*)wl;
  const double e0 = ext[0], e1 = ext[1];
  return pmap(*n, [=] __device__(int64_t i) {
    Ue[i] = e0 * U[i] + e1 * Ul[i];
    Ve[i] = e0 * V[i] + e1 * Vl[i];
    We[i] = e0 * W[i] + e1 * Wl[i];
  });
}

// makeabf + makebdf + lag update, per component c (three calls):
//   F = ext1 f + ext2 flag + (B rho/dt)(bd1 u + bd2 ulag);  flag = f;  ulag = u
// coef = {ext1, ext2, rho/dt * bd1, rho/dt * bd2}.
int nekoamd_pnpn_makef(void* F, void* flag, void* ulag, const void* f,
                       const void* u, const void* B, const double* coef,
                       const int64_t* n) {
  NK_PN(n);
  NK_REQUIRE(coef, "null coefficients");
  double *Fo = (double*)F, *Fl = (double*)flag, *Ul = (double*)ulag;
  const double *Fn = (const double*)f, *U = (const double*)u,
               *Bm = (const double*)B;
  const double e0 = coef[0], e1 = coef[1], b1 = coef[2], b2 = coef[3];
  return pmap(*n, [=] __device__(int64_t i) {
    const double fi = Fn[i], ui = U[i], b = Bm[i];
    Fo[i] = e0 * fi + e1 * Fl[i] + b * (b1 * ui + b2 * Ul[i]);
    Fl[i] = fi;
    Ul[i] = ui;
  });
}

// The three components of nekoamd_pnpn_makef in one pass (B read once: 176
// B per point instead of 192); each component gets the same operations in
// the same order as its own call.
int nekoamd_pnpn_makef_vector(void* F1, void* F2, void* F3, void* flag1,
                              void* flag2, void* flag3, void* ulag1,
                              void* ulag2, void* ulag3, const void* f1,
                              const void* f2, const void* f3, const void* u1,
                              const void* u2, const void* u3, const void* B,
                              const double* coef, const int64_t* n) {
  NK_PN(n);
  NK_REQUIRE(coef, "null coefficients");
  NK_REQUIRE(*n == 0 || (F1 && F2 && F3 && flag1 && flag2 && flag3 && ulag1 &&
                         ulag2 && ulag3 && f1 && f2 && f3 && u1 && u2 && u3 && B),
             "null field");
  struct Comp {
    double *Fo, *Fl, *Ul;
    const double *Fn, *U;
  };
  const Comp c0{(double*)F1, (double*)flag1, (double*)ulag1, (const double*)f1,
                (const double*)u1};
  const Comp c1{(double*)F2, (double*)flag2, (double*)ulag2, (const double*)f2,
                (const double*)u2};
  const Comp c2{(double*)F3, (double*)flag3, (double*)ulag3, (const double*)f3,
                (const double*)u3};
  const double* Bm = (const double*)B;
  const double e0 = coef[0], e1 = coef[1], b1 = coef[2], b2 = coef[3];
  return pmap(*n, [=] __device__(int64_t i) {
    const double b = Bm[i];
    auto one = [&](const Comp& c) {
      const double fi = c.Fn[i], ui = c.U[i];
      c.Fo[i] = e0 * fi + e1 * c.Fl[i] + b * (b1 * ui + b2 * c.Ul[i]);
      c.Fl[i] = fi;
      c.Ul[i] = ui;
    };
    one(c0);
    one(c1);
    one(c2);
  });
}

// time_order 3 (BDF3/EXT3): the extrapolation with the second lags,
//   ue = ext1 u + ext2 ulag + ext3 ulag2
int nekoamd_pnpn_extrapolate3(void* ue, void* ve, void* we, const void* u,
                              const void* v, const void* w, const void* ul,
                              const void* vl, const void* wl, const void* ul2,
                              const void* vl2, const void* wl2,
                              const double* ext, const int64_t* n) {
  NK_PN(n);
  NK_REQUIRE(ext, "null coefficients");
  double *Ue = (double*)ue, *Ve = (double*)ve, *We = (double*)we;
  const double *U = (const double*)u, *V = (const double*)v,
               *W = (const double*)w, *Ul = (const double*)ul,
               *Vl = (const double*)vl, *Wl = (const double*)wl,
               *Ul2 = (const double*)ul2, *Vl2 = (const double*)vl2,
               *Wl2 = (const double*)wl2;
  const double e0 = ext[0], e1 = ext[1], e2 = ext[2];
  return pmap(*n, [=] __device__(int64_t i) {
    Ue[i] = e0 * U[i] + e1 * Ul[i] + e2 * Ul2[i];
    Ve[i] = e0 * V[i] + e1 * Vl[i] + e2 * Vl2[i];
    We[i] = e0 * W[i] + e1 * Wl[i] + e2 * Wl2[i];
  });
}

// time_order 3: makeabf + makebdf + lag update with the second lags,
//   F = ext1 f + ext2 flag + ext3 flag2 + (B rho/dt)(bd1 u + bd2 ulag + bd3 ulag2);
//   flag2 = flag; flag = f; ulag2 = ulag; ulag = u
// coef = {ext1, ext2, ext3, rho/dt * bd1, rho/dt * bd2, rho/dt * bd3}.
int nekoamd_pnpn_makef3(void* F, void* flag, void* flag2, void* ulag,
                        void* ulag2, const void* f, const void* u,
                        const void* B, const double* coef, const int64_t* n) {
  NK_PN(n);
  NK_REQUIRE(coef, "null coefficients");
  double *Fo = (double*)F, *Fl = (double*)flag, *Fl2 = (double*)flag2,
         *Ul = (double*)ulag, *Ul2 = (double*)ulag2;
  const double *Fn = (const double*)f, *U = (const double*)u,
               *Bm = (const double*)B;
  const double e0 = coef[0], e1 = coef[1], e2 = coef[2], b1 = coef[3],
               b2 = coef[4], b3 = coef[5];
  return pmap(*n, [=] __device__(int64_t i) {
    const double fi = Fn[i], ui = U[i], b = Bm[i], fl = Fl[i], ul = Ul[i];
    const double fl2 = Fl2[i], ul2 = Ul2[i];
    Fo[i] = e0 * fi + e1 * fl + e2 * fl2 + b * (b1 * ui + b2 * ul + b3 * ul2);
    Fl2[i] = fl;
    Fl[i] = fi;
    Ul2[i] = ul;
    Ul[i] = ui;
  });
}

// Pressure forcing, per component: t = F/rho - mu B c - chi B ue
int nekoamd_pnpn_prs_force(void* t, const void* F, const void* c,
                           const void* ue, const void* chi, const void* B,
                           const double* mu, const double* rho,
                           const int64_t* n) {
  NK_PN(n);
  NK_REQUIRE(mu && rho, "null scalar");
  double* T = (double*)t;
  const double *Fv = (const double*)F, *Cv = (const double*)c,
               *Ue = (const double*)ue, *X = (const double*)chi,
               *Bm = (const double*)B;
  const double m = *mu, r = *rho;
  return pmap(*n, [=] __device__(int64_t i) {
    const double b = Bm[i];
    T[i] = Fv[i] / r - m * b * Cv[i] - X[i] * b * Ue[i];
  });
}

// The three components of nekoamd_pnpn_prs_force in one pass: chi and B
// read once for the three (112 B per point instead of 144); each component
// gets the same operations in the same order as its own call.
int nekoamd_pnpn_prs_force_vector(void* t1, void* t2, void* t3, const void* F1,
                                  const void* F2, const void* F3,
                                  const void* c1, const void* c2,
                                  const void* c3, const void* ue1,
                                  const void* ue2, const void* ue3,
                                  const void* chi, const void* B,
                                  const double* mu, const double* rho,
                                  const int64_t* n) {
  NK_PN(n);
  NK_REQUIRE(mu && rho, "null scalar");
  NK_REQUIRE(*n == 0 || (t1 && t2 && t3 && F1 && F2 && F3 && c1 && c2 && c3 &&
                         ue1 && ue2 && ue3 && chi && B),
             "null field");
  double *T1 = (double*)t1, *T2 = (double*)t2, *T3 = (double*)t3;
  const double *Fa = (const double*)F1, *Fb = (const double*)F2,
               *Fc = (const double*)F3, *Ca = (const double*)c1,
               *Cb = (const double*)c2, *Cc = (const double*)c3,
               *Ua = (const double*)ue1, *Ub = (const double*)ue2,
               *Uc = (const double*)ue3, *X = (const double*)chi,
               *Bm = (const double*)B;
  const double m = *mu, r = *rho;
  return pmap(*n, [=] __device__(int64_t i) {
    const double b = Bm[i], x = X[i];
    T1[i] = Fa[i] / r - m * b * Ca[i] - x * b * Ua[i];
    T2[i] = Fb[i] / r - m * b * Cb[i] - x * b * Ub[i];
    T3[i] = Fc[i] / r - m * b * Cc[i] - x * b * Uc[i];
  });
}

// a = a - b - s c  (pressure residual: cdtp - A p - (rho bd0/dt) flux;
// velocity residual: F - opgrad - H u with s = 0 and c unused when null).
int nekoamd_sub2s(void* a, const void* b, const void* c, const double* s,
                  const int64_t* n) {
  NK_PN(n);
  NK_REQUIRE(s, "null scalar");
  double* A = (double*)a;
  const double *Bv = (const double*)b, *Cv = (const double*)c;
  const double S = *s;
  if (!c)
    return pmap(*n, [=] __device__(int64_t i) { A[i] = A[i] - Bv[i]; });
  return pmap(*n,
              [=] __device__(int64_t i) { A[i] = A[i] - Bv[i] - S * Cv[i]; });
}

// r = F - g - r   (velocity residual with r = H u on entry)
int nekoamd_pnpn_vel_res(void* r, const void* F, const void* g,
                         const int64_t* n) {
  NK_PN(n);
  double* R = (double*)r;
  const double *Fv = (const double*)F, *G = (const double*)g;
  return pmap(*n, [=] __device__(int64_t i) { R[i] = Fv[i] - G[i] - R[i]; });
}

// out[l] = U[0,1) from SplitMix64(seed, gid(l)) of the box lattice id, so the
// field is continuous (one value per global point).
int nekoamd_box_rng_uniform(void* out, const int* nx, const int* ny,
                            const int* nz, const int* lx,
                            const uint64_t* seed) {
  NK_INIT();
  NK_REQUIRE(out && nx && ny && nz && lx && seed, "null argument");
  const int NX = *nx, NY = *ny, L = *lx, P = L - 1;
  const int64_t n = (int64_t)NX * NY * *nz * L * L * L;
  const int64_t sx = (int64_t)NX * P + 1, sy = (int64_t)NY * P + 1;
  const uint64_t S = *seed;
  double* O = (double*)out;
  return pmap(n, [=] __device__(int64_t l) {
    const int64_t np = (int64_t)L * L * L;
    const int64_t e = l / np, q = l % np;
    const int i = q % L, j = (q / L) % L, k = q / (L * L);
    const int64_t ex = e % NX, ey = (e / NX) % NY, ez = e / ((int64_t)NX * NY);
    const int64_t gid = (ex * P + i) + sx * ((ey * P + j) + sy * (ez * P + k));
    uint64_t z = S + ((uint64_t)gid + 1u) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    O[l] = (double)(z >> 11) * 0x1.0p-53;
  });
}

}  // extern "C"
