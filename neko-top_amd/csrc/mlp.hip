// mlp.hip -- the element-local network of the learned solver aid: BF16
// MFMA forward and backward, Adam, harvesting of (residual, solution) pairs
// from the solver, training and inference on gfx950.
//
// The reference has no learned component; its initial-guess accelerator is
// the projection of sources/adjoint/adjoint_pnpn.f90:102-105, :729-741
// (projection.hip).  This network adds one more direction to the initial
// guess of the pressure solve:
//   q = mask(mult QQ^T N(b)),  gamma = <q,b>/<q,Aq>,  x0 += gamma q,
//   b -= gamma A q
// where N maps, per element, the 512 values of the residual b and the 6
// face-neighbour means of b (all divided by the RMS of b) to 512 values:
//   518 -> 256 (SiLU) -> 256 (SiLU) -> 512.
// The Galerkin scaling makes the guess safe: it never increases the A-norm
// of the error, and gamma = 0 (no guess) when <q,Aq> <= 0.
//
// GEMMs: one 256-thread kernel, 128 x 128 x 32 tiles staged through LDS,
// v_mfma_f32_32x32x16_bf16 (four waves, 64 x 64 each), FP32 accumulation.
// Three operand layouts (NT forward, NN data gradient, TN weight gradient
// with split-K) are converted to one K-contiguous LDS layout on load; the
// bias, SiLU, SiLU' and BF16 rounding are fused into the epilogues.
// BF16 is stored as uint16_t and rounded to nearest-even by bit arithmetic.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "internal.h"
#include "nekoamd.h"
#include "reduce.h"

namespace nekoamd {

constexpr int kNP = 512;    // GLL values per element (lx = 8)
constexpr int kNIN = 518;   // inputs
constexpr int kKIN = 544;   // inputs padded to the K tile
constexpr int kHID = 256;
constexpr int kBM = 128, kBN = 128, kBK = 32, kLdsK = kBK + 8;

// padded parameter layout (FP32): W1[256][544] b1 W2[256][256] b2
// W3[512][256] b3
constexpr int64_t kOffW1 = 0;
constexpr int64_t kOffB1 = kOffW1 + (int64_t)kHID * kKIN;
constexpr int64_t kOffW2 = kOffB1 + kHID;
constexpr int64_t kOffB2 = kOffW2 + (int64_t)kHID * kHID;
constexpr int64_t kOffW3 = kOffB2 + kHID;
constexpr int64_t kOffB3 = kOffW3 + (int64_t)kNP * kHID;
constexpr int64_t kNParPad = kOffB3 + kNP;
static_assert(NEKOAMD_MLP_NPARAM ==
                  kHID * kNIN + kHID + kHID * kHID + kHID + kNP * kHID + kNP,
              "parameter count");

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) {
  return __uint_as_float((uint32_t)h << 16);
}
__device__ __forceinline__ float silu(float z) {
  return z / (1.0f + __expf(-z));
}
__device__ __forceinline__ float dsilu(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

enum Layout { kNT = 0, kNN = 1, kTN = 2 };
enum Epi { kFwdHidden = 0, kFwdOut = 1, kBwdData = 2, kWgrad = 3 };

struct EpiArgs {
  const float* bias;  // FwdHidden / FwdOut
  float* zf;          // FwdHidden: Z (may be null); BwdData: dZ fp32; Wgrad: partials
  uint16_t* hb;       // FwdHidden: H; BwdData: dZ bf16
  const float* z;     // BwdData: the forward pre-activation
  float* yf;          // FwdOut (float output)
  double* yd;         // FwdOut (double output, scaled by *yscale)
  const double* yscale;
  int ldc;
};

// C[M,N] = sum_k opA[m,k] opB[n,k]:
//   NT: A[m*lda+k], B[n*ldb+k];  NN: A[m*lda+k], B[k*ldb+n];
//   TN: A[k*lda+m], B[k*ldb+n].
// K (NT: also the contiguous dim), and the contiguous dims of NN / TN
// operands, must be multiples of 8; other dims are guarded.  Split-K
// (gridDim.z > 1, Wgrad only): block z covers k in [z*kz, (z+1)*kz).
// BM x BN output tile per 256-thread workgroup, four waves in a 2 x 2
// arrangement, each wave (BM/2) x (BN/2) as (BM/64) x (BN/64) MFMA tiles of
// 32 x 32.  128 x 128 for the training GEMMs; 64 x 64 for the forward ones,
// whose M = elements is tall and K <= 544 small: four times the workgroups
// (8 waves per SIMD at 32^3) hide the global-load latency of each k-step
// (inference at 32^3: 116 -> 77 us; prefetching the next k-tile into
// registers on top of that measured no gain).
template <int LAYOUT, int EPI, int BM = kBM, int BN = kBN>
__global__ void __launch_bounds__(256)
    gemm_bf16_k(int M, int N, int K, int kz, const uint16_t* __restrict__ A,
                int lda, const uint16_t* __restrict__ B, int ldb, EpiArgs ep) {
  static_assert(BM % 64 == 0 && BN % 64 == 0, "tile must be a multiple of 64");
  constexpr int TI = BM / 64, TJ = BN / 64;
  __shared__ __attribute__((aligned(16))) uint16_t As[BM * kLdsK];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[BN * kLdsK];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const int kbeg = blockIdx.z * kz;
  const int kend = min(K, kbeg + kz);

  f32x16 acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j)
#pragma unroll
      for (int a = 0; a < 16; ++a) acc[i][j][a] = 0.0f;

  for (int k0 = kbeg; k0 < kend; k0 += kBK) {
    // ---- stage A tile into As[m][k] ----
#pragma unroll
    for (int c2 = 0; c2 < BM / 64; ++c2) {
      const int c = t + c2 * 256;
      if (LAYOUT == kTN) {
        const int kk = c / (BM / 8), mc = (c % (BM / 8)) * 8;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (k0 + kk < kend && m0 + mc < M)
          v = *(const uint4*)(A + (int64_t)(k0 + kk) * lda + m0 + mc);
        const uint16_t* h = (const uint16_t*)&v;
#pragma unroll
        for (int j = 0; j < 8; ++j) As[(mc + j) * kLdsK + kk] = h[j];
      } else {
        const int r = c >> 2, kc = (c & 3) * 8;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (m0 + r < M && k0 + kc < kend)
          v = *(const uint4*)(A + (int64_t)(m0 + r) * lda + k0 + kc);
        *(uint4*)(As + r * kLdsK + kc) = v;
      }
    }
    // ---- stage B tile into Bs[n][k] ----
#pragma unroll
    for (int c2 = 0; c2 < BN / 64; ++c2) {
      const int c = t + c2 * 256;
      if (LAYOUT == kNT) {
        const int r = c >> 2, kc = (c & 3) * 8;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (n0 + r < N && k0 + kc < kend)
          v = *(const uint4*)(B + (int64_t)(n0 + r) * ldb + k0 + kc);
        *(uint4*)(Bs + r * kLdsK + kc) = v;
      } else {
        const int kk = c / (BN / 8), nc = (c % (BN / 8)) * 8;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (k0 + kk < kend && n0 + nc < N)
          v = *(const uint4*)(B + (int64_t)(k0 + kk) * ldb + n0 + nc);
        const uint16_t* h = (const uint16_t*)&v;
#pragma unroll
        for (int j = 0; j < 8; ++j) Bs[(nc + j) * kLdsK + kk] = h[j];
      }
    }
    __syncthreads();
    // ---- MFMA: each wave TI x TJ tiles of 32 x 32, two k-steps ----
#pragma unroll
    for (int ks = 0; ks < kBK; ks += 16) {
      const int kq = ks + (lane >> 5) * 8;
      bf16x8 a[TI], b[TJ];
#pragma unroll
      for (int i = 0; i < TI; ++i) {
        const int row = wm * (BM / 2) + i * 32 + (lane & 31);
        const uint4 v = *(const uint4*)(As + row * kLdsK + kq);
        a[i] = __builtin_bit_cast(bf16x8, v);
      }
#pragma unroll
      for (int j = 0; j < TJ; ++j) {
        const int col = wn * (BN / 2) + j * 32 + (lane & 31);
        const uint4 v = *(const uint4*)(Bs + col * kLdsK + kq);
        b[j] = __builtin_bit_cast(bf16x8, v);
      }
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j],
                                                              acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }

  // ---- epilogue: D[row][col], col = lane % 32, row = 8(a/4) + 4(lane/32) + a%4
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j)
#pragma unroll
      for (int a = 0; a < 16; ++a) {
        const int row = m0 + wm * (BM / 2) + i * 32 + 8 * (a >> 2) +
                        4 * (lane >> 5) + (a & 3);
        const int col = n0 + wn * (BN / 2) + j * 32 + (lane & 31);
        if (row >= M || col >= N) continue;
        const int64_t o = (int64_t)row * ep.ldc + col;
        const float v = acc[i][j][a];
        if constexpr (EPI == kFwdHidden) {
          const float zz = v + ep.bias[col];
          if (ep.zf) ep.zf[o] = zz;
          ep.hb[o] = f2bf(silu(zz));
        } else if constexpr (EPI == kFwdOut) {
          const float y = v + ep.bias[col];
          if (ep.yf) ep.yf[o] = y;
          if (ep.yd) ep.yd[o] = (double)y * ep.yscale[0];
        } else if constexpr (EPI == kBwdData) {
          const float dz = v * dsilu(ep.z[o]);
          ep.zf[o] = dz;
          ep.hb[o] = f2bf(dz);
        } else {
          ep.zf[(int64_t)blockIdx.z * M * ep.ldc + o] = v;
        }
      }
}

// ---------------------------------------------------------------------------
// features, targets, scales

// part[b] = partial sum of u^2 over block b's share (fixed order).
__global__ void __launch_bounds__(kRedBlock)
    mlp_sumsq_k(int64_t n, const double* __restrict__ u,
                double* __restrict__ part) {
  __shared__ double sh[kRedBlock / 64];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRedBlock)
    s += u[i] * u[i];
  s = block_sum<kRedBlock>(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// sc[0] = RMS = sqrt(sum/n), sc[1] = 1/RMS (1 when the field is zero).
__global__ void __launch_bounds__(kFinBlock)
    mlp_rms_fin_k(const double* __restrict__ part, int np, double n,
                  double* __restrict__ sc) {
  __shared__ double sh[kFinBlock / 64];
  const double s = sum_partials<kFinBlock>(part, np, sh);
  if (threadIdx.x == 0) {
    const double r = sqrt(s / n);
    sc[0] = r > 0.0 ? r : 1.0;
    sc[1] = r > 0.0 ? 1.0 / r : 1.0;
  }
}

// mean[e] = mean of the element's 512 values of u.
__global__ void __launch_bounds__(256)
    mlp_mean_k(const double* __restrict__ u, double* __restrict__ mean) {
  __shared__ double sh[256 / 64];
  const int64_t e = blockIdx.x;
  double s = u[e * kNP + threadIdx.x] + u[e * kNP + 256 + threadIdx.x];
  s = block_sum<256>(s, sh);
  if (threadIdx.x == 0) mean[e] = s * (1.0 / kNP);
}

// X[e][0..512) = bf16(u_e / s), X[e][512+k] = bf16(mean[nb] / s), rest 0.
__global__ void __launch_bounds__(64)
    mlp_features_k(const double* __restrict__ u,
                   const double* __restrict__ mean,
                   const int32_t* __restrict__ nbr,
                   const double* __restrict__ sc, uint16_t* __restrict__ X) {
  const int64_t e = blockIdx.x;
  const int t = threadIdx.x;
  const double inv = sc[1];
  uint16_t* xe = X + e * kKIN;
  uint16_t h[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) h[j] = f2bf((float)(u[e * kNP + t * 8 + j] * inv));
  uint4 hv;
  memcpy(&hv, h, sizeof(hv));
  *(uint4*)(xe + t * 8) = hv;
  if (t < 32) {
    const int c = kNP + t;  // 512 .. 543
    uint16_t v = 0;
    if (t < 6) {
      const int j = nbr[e * 6 + t];
      v = j >= 0 ? f2bf((float)(mean[j] * inv)) : (uint16_t)0;
    }
    xe[c] = v;
  }
}

// T[i] = x[i] / s (FP32)
__global__ void __launch_bounds__(256)
    mlp_target_k(int64_t n, const double* __restrict__ x,
                 const double* __restrict__ sc, float* __restrict__ T) {
  const double inv = sc[1];
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256)
    T[i] = (float)(x[i] * inv);
}

// dY = 2 (Y - T inv_ysc) / (M 512) as FP32 and BF16; part[b] = partial of
// (Y - T inv_ysc)^2.
__global__ void __launch_bounds__(256)
    mlp_dloss_k(int64_t n, const float* __restrict__ Y,
                const float* __restrict__ T, const double* __restrict__ ysc,
                float* __restrict__ dzf, uint16_t* __restrict__ dzb,
                double* __restrict__ part) {
  __shared__ double sh[256 / 64];
  const float inv = (float)ysc[1];
  const float g = 2.0f / (float)n;
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    const float d = Y[i] - T[i] * inv;
    s += (double)d * (double)d;
    const float dz = g * d;
    dzf[i] = dz;
    dzb[i] = f2bf(dz);
  }
  s = block_sum<256>(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Column sums db[c] = sum_r dz[r][c] in two fixed-order levels:
// part[chunk][c] = sum of kColRows consecutive rows, then db[c] = sum of the
// chunks.  kColRows rows per thread, unrolled so the loads overlap.
constexpr int kColRows = 64;

__global__ void __launch_bounds__(256)
    mlp_colsum_part_k(int M, int N, const float* __restrict__ dz,
                      float* __restrict__ part) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= N) return;
  const int r0 = blockIdx.y * kColRows;
  const int r1 = min(M, r0 + kColRows);
  float s = 0.0f;
#pragma unroll 16
  for (int r = r0; r < r1; ++r) s += dz[(int64_t)r * N + c];
  part[(int64_t)blockIdx.y * N + c] = s;
}

__global__ void __launch_bounds__(256)
    mlp_colsum_fin_k(int nchunk, int N, const float* __restrict__ part,
                     float* __restrict__ db) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= N) return;
  float s = 0.0f;
  for (int q = 0; q < nchunk; ++q) s += part[(int64_t)q * N + c];
  db[c] = s;
}

// g[i] = sum_z ws[z][i] for the first ncols of every row of width ld.
__global__ void __launch_bounds__(256)
    mlp_splitk_sum_k(int64_t rows, int ld, int ncols, int nz,
                     const float* __restrict__ ws, float* __restrict__ g) {
  const int64_t n = rows * ld;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    const int col = (int)(i % ld);
    float s = 0.0f;
    if (col < ncols)
      for (int z = 0; z < nz; ++z) s += ws[(int64_t)z * n + i];
    g[i] = s;
  }
}

// Adam on the padded FP32 parameters; refreshes the BF16 weight copies.
// Padding entries have zero gradient and stay zero.
__global__ void __launch_bounds__(256)
    mlp_adam_k(float* __restrict__ th, const float* __restrict__ g,
               float* __restrict__ m, float* __restrict__ v, float lr,
               float b1, float b2, float c1, float c2, float eps,
               uint16_t* __restrict__ w1b, uint16_t* __restrict__ w2b,
               uint16_t* __restrict__ w3b) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < kNParPad;
       i += (int64_t)gridDim.x * 256) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float t = th[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    th[i] = t;
    if (i < kOffB1)
      w1b[i - kOffW1] = f2bf(t);
    else if (i >= kOffW2 && i < kOffB2)
      w2b[i - kOffW2] = f2bf(t);
    else if (i >= kOffW3 && i < kOffB3)
      w3b[i - kOffW3] = f2bf(t);
  }
}

__global__ void __launch_bounds__(256)
    mlp_tobf16_k(int64_t n, const float* __restrict__ a,
                 uint16_t* __restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256)
    b[i] = f2bf(a[i]);
}

// out (unpadded, 330,240) <- in (padded)
__global__ void __launch_bounds__(256)
    mlp_unpad_k(const float* __restrict__ in, float* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
       i < NEKOAMD_MLP_NPARAM; i += (int64_t)gridDim.x * 256) {
    int64_t src;
    if (i < (int64_t)kHID * kNIN)
      src = (i / kNIN) * kKIN + (i % kNIN);
    else
      src = kOffB1 + (i - (int64_t)kHID * kNIN);
    out[i] = in[src];
  }
}

__global__ void __launch_bounds__(256)
    mlp_axpy_k(int64_t n, double* __restrict__ y, const double* __restrict__ x,
               const double* __restrict__ a) {
  const double g = a[0];
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256)
    y[i] += g * x[i];
}

__global__ void __launch_bounds__(256)
    mlp_mulmask_k(int64_t n, double* __restrict__ q,
                  const double* __restrict__ mult) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256)
    q[i] *= mult[i];
}

// part = partials of {<q,b>_mult, <q,aq>_mult, <b,b>_mult}
__global__ void __launch_bounds__(kRedBlock)
    mlp_dots_k(int64_t n, const double* __restrict__ q,
               const double* __restrict__ b, const double* __restrict__ aq,
               const double* __restrict__ mult, double* __restrict__ part) {
  __shared__ double sh[kRedBlock / 64];
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRedBlock) {
    const double m = mult[i];
    s0 += m * q[i] * b[i];
    s1 += m * q[i] * aq[i];
    s2 += m * b[i] * b[i];
  }
  s0 = block_sum<kRedBlock>(s0, sh);
  s1 = block_sum<kRedBlock>(s1, sh);
  s2 = block_sum<kRedBlock>(s2, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s0;
    part[kRedGrid + blockIdx.x] = s1;
    part[2 * kRedGrid + blockIdx.x] = s2;
  }
}

// sc[0..2] = the three sums; sc[3] = gamma = <q,b>/<q,aq> (0 unless
// <q,aq> > 0 and the result is finite).
__global__ void __launch_bounds__(kFinBlock)
    mlp_gamma_k(const double* __restrict__ part, int np,
                double* __restrict__ sc) {
  __shared__ double sh[kFinBlock / 64];
  const double qb = sum_partials<kFinBlock>(part, np, sh);
  const double qaq = sum_partials<kFinBlock>(part + kRedGrid, np, sh);
  const double bb = sum_partials<kFinBlock>(part + 2 * kRedGrid, np, sh);
  if (threadIdx.x == 0) {
    sc[0] = qb;
    sc[1] = qaq;
    sc[2] = bb;
    const double g = qaq > 0.0 ? qb / qaq : 0.0;
    sc[3] = isfinite(g) ? g : 0.0;
    sc[4] = -sc[3];
  }
}

// part[b] = partial sum of T^2 (FP32 values, FP64 sums)
__global__ void __launch_bounds__(kRedBlock)
    mlp_sumsq_f_k(int64_t n, const float* __restrict__ u,
                  double* __restrict__ part) {
  __shared__ double sh[kRedBlock / 64];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRedBlock)
    s += (double)u[i] * (double)u[i];
  s = block_sum<kRedBlock>(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// acc[0] += sum of the np loss partials (one block, fixed order).
__global__ void __launch_bounds__(kFinBlock)
    mlp_loss_acc_k(const double* __restrict__ part, int np,
                   double* __restrict__ acc) {
  __shared__ double sh[kFinBlock / 64];
  const double s = sum_partials<kFinBlock>(part, np, sh);
  if (threadIdx.x == 0) acc[0] += s;
}

// ---------------------------------------------------------------------------
struct MlpState {
  int nelv = 0, nx = 0, ny = 0, nz = 0;
  int64_t n = 0;
  int cap = 0, npairs = 0, next = 0;
  int64_t adam_t = 0;
  int32_t* nbr = nullptr;
  float *th = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  uint16_t *w1b = nullptr, *w2b = nullptr, *w3b = nullptr;
  // training set: X [cap*nelv][544] bf16, T [cap*nelv][512] f32
  uint16_t* Xd = nullptr;
  float* Td = nullptr;
  // per-batch work (rows = max(batch, nelv))
  int rows = 0;
  uint16_t *xb = nullptr, *h1 = nullptr, *h2 = nullptr, *dzb = nullptr;
  float *z1 = nullptr, *z2 = nullptr, *y = nullptr, *dzf = nullptr;
  float* ws = nullptr;  // split-K partials
  int64_t ws_len = 0;
  double* mean = nullptr;  // nelv
  double* sc = nullptr;    // scalars: [0..1] rms/inv of the input field,
                           // [2..3] target scale, [8..12] guess dots/gamma
  double* part = nullptr;  // 3 * kRedGrid
  double* q = nullptr;     // n: the guess direction
  double* aq = nullptr;    // n
  double* bsave = nullptr; // n: residual kept for harvesting
  bool applied = false, saved = false;
  double* host = nullptr;  // pinned, 16
};

static MlpState* as_mlp(void* p) { return (MlpState*)p; }

// K range per split (a multiple of the K tile) and the number of splits.
static int split_kz(int K, int splits) {
  if (splits <= 1) return K;
  return ((K + splits - 1) / splits + kBK - 1) / kBK * kBK;
}
static int split_nz(int K, int splits) {
  const int kz = split_kz(K, splits);
  return kz > 0 ? (K + kz - 1) / kz : 1;
}

template <int LAYOUT, int EPI, int BM = kBM, int BN = kBN>
static int gemm(int M, int N, int K, const uint16_t* A, int lda,
                const uint16_t* B, int ldb, const EpiArgs& ep, int splits = 1) {
  if (M <= 0 || N <= 0) return kOk;
  const int kz = split_kz(K, splits);
  const int nz = split_nz(K, splits);
  dim3 grid((M + BM - 1) / BM, (N + BN - 1) / BN, nz);
  hipLaunchKernelGGL((gemm_bf16_k<LAYOUT, EPI, BM, BN>), grid, dim3(256), 0,
                     stream(), M, N, K, kz, A, lda, B, ldb, ep);
  NK_LAUNCHED();
  return kOk;
}

// db = column sums of dz[M][N]; uses the split-K workspace as scratch.
static int colsum(MlpState* s, int M, int N, const float* dz, float* db) {
  const int nchunk = (M + kColRows - 1) / kColRows;
  hipLaunchKernelGGL(mlp_colsum_part_k, dim3((N + 255) / 256, nchunk),
                     dim3(256), 0, stream(), M, N, dz, s->ws);
  NK_LAUNCHED();
  hipLaunchKernelGGL(mlp_colsum_fin_k, dim3((N + 255) / 256), dim3(256), 0,
                     stream(), nchunk, N, s->ws, db);
  NK_LAUNCHED();
  return kOk;
}

static int nsplits(int K) {
  int z = K / 512;
  return z < 1 ? 1 : (z > 32 ? 32 : z);
}

// Forward over M rows of X into the state's work buffers.
static int forward(MlpState* s, const uint16_t* X, int M, bool keep_z,
                   float* yf, double* yd, const double* yscale) {
  EpiArgs e{};
  e.bias = s->th + kOffB1;
  e.zf = keep_z ? s->z1 : nullptr;
  e.hb = s->h1;
  e.ldc = kHID;
  int rc = gemm<kNT, kFwdHidden, 64, 64>(M, kHID, kKIN, X, kKIN, s->w1b, kKIN, e);
  if (rc) return rc;
  e.bias = s->th + kOffB2;
  e.zf = keep_z ? s->z2 : nullptr;
  e.hb = s->h2;
  rc = gemm<kNT, kFwdHidden, 64, 64>(M, kHID, kHID, s->h1, kHID, s->w2b, kHID, e);
  if (rc) return rc;
  EpiArgs o{};
  o.bias = s->th + kOffB3;
  o.yf = yf;
  o.yd = yd;
  o.yscale = yscale;
  o.ldc = kNP;
  return gemm<kNT, kFwdOut, 64, 64>(M, kNP, kHID, s->h2, kHID, s->w3b, kHID, o);
}

// Gradient of the batch (X, T) into s->g (padded layout); loss partials in
// s->part[0..np).  inv_ysc: device scalar the targets are multiplied by.
static int backward(MlpState* s, const uint16_t* X, const float* T, int M,
                    const double* ysc, int* np_out) {
  int rc = forward(s, X, M, true, s->y, nullptr, nullptr);
  if (rc) return rc;
  const int64_t ny = (int64_t)M * kNP;
  const int gl = nk_grid(ny, 256, kRedGrid);
  hipLaunchKernelGGL(mlp_dloss_k, dim3(gl), dim3(256), 0, stream(), ny, s->y,
                     T, ysc, s->dzf, s->dzb, s->part);
  NK_LAUNCHED();
  *np_out = gl;
  // layer 3
  rc = colsum(s, M, kNP, s->dzf, s->g + kOffB3);
  if (rc) return rc;
  const int sp = nsplits(M);
  EpiArgs w{};
  w.zf = s->ws;
  w.ldc = kHID;
  rc = gemm<kTN, kWgrad>(kNP, kHID, M, s->dzb, kNP, s->h2, kHID, w, sp);
  if (rc) return rc;
  const int nz = split_nz(M, sp);
  hipLaunchKernelGGL(mlp_splitk_sum_k, dim3(nk_grid((int64_t)kNP * kHID, 256)),
                     dim3(256), 0, stream(), (int64_t)kNP, kHID, kHID, nz,
                     s->ws, s->g + kOffW3);
  NK_LAUNCHED();
  // dZ2 = (dZ3 W3) * silu'(Z2): dZ3 lives in dz*[0, M*512), dZ2 goes
  // behind it; dZ1 then reuses the front of the buffers.
  uint16_t* dz2b = s->dzb + ny;
  float* dz2f = s->dzf + ny;
  EpiArgs d{};
  d.z = s->z2;
  d.zf = dz2f;
  d.hb = dz2b;
  d.ldc = kHID;
  rc = gemm<kNN, kBwdData>(M, kHID, kNP, s->dzb, kNP, s->w3b, kHID, d);
  if (rc) return rc;
  rc = colsum(s, M, kHID, dz2f, s->g + kOffB2);
  if (rc) return rc;
  rc = gemm<kTN, kWgrad>(kHID, kHID, M, dz2b, kHID, s->h1, kHID, w, sp);
  if (rc) return rc;
  hipLaunchKernelGGL(mlp_splitk_sum_k, dim3(nk_grid((int64_t)kHID * kHID, 256)),
                     dim3(256), 0, stream(), (int64_t)kHID, kHID, kHID, nz,
                     s->ws, s->g + kOffW2);
  NK_LAUNCHED();
  // dZ1 = (dZ2 W2) * silu'(Z1) into the first part of the buffers
  d.z = s->z1;
  d.zf = s->dzf;
  d.hb = s->dzb;
  rc = gemm<kNN, kBwdData>(M, kHID, kHID, dz2b, kHID, s->w2b, kHID, d);
  if (rc) return rc;
  rc = colsum(s, M, kHID, s->dzf, s->g + kOffB1);
  if (rc) return rc;
  w.ldc = kKIN;
  rc = gemm<kTN, kWgrad>(kHID, kKIN, M, s->dzb, kHID, X, kKIN, w, sp);
  if (rc) return rc;
  hipLaunchKernelGGL(mlp_splitk_sum_k, dim3(nk_grid((int64_t)kHID * kKIN, 256)),
                     dim3(256), 0, stream(), (int64_t)kHID, kKIN, kNIN, nz,
                     s->ws, s->g + kOffW1);
  NK_LAUNCHED();
  return kOk;
}

static int rms(MlpState* s, const double* u, double* out) {
  const int g = nk_grid(s->n, kRedBlock, kRedGrid);
  hipLaunchKernelGGL(mlp_sumsq_k, dim3(g), dim3(kRedBlock), 0, stream(), s->n,
                     u, s->part);
  NK_LAUNCHED();
  hipLaunchKernelGGL(mlp_rms_fin_k, dim3(1), dim3(kFinBlock), 0, stream(),
                     s->part, g, (double)s->n, out);
  NK_LAUNCHED();
  return kOk;
}

// X rows of one field (nelv rows) from u with scale sc[0..1].
static int make_features(MlpState* s, const double* u, const double* sc,
                         uint16_t* X) {
  hipLaunchKernelGGL(mlp_mean_k, dim3(s->nelv), dim3(256), 0, stream(), u,
                     s->mean);
  NK_LAUNCHED();
  hipLaunchKernelGGL(mlp_features_k, dim3(s->nelv), dim3(64), 0, stream(), u,
                     s->mean, s->nbr, sc, X);
  NK_LAUNCHED();
  return kOk;
}

template <typename T>
static T* dalloc(int64_t n, hipError_t& e) {
  T* p = nullptr;
  if (e == hipSuccess) e = hipMalloc((void**)&p, sizeof(T) * (n > 0 ? n : 1));
  return p;
}

static int refresh_bf16(MlpState* s) {
  hipLaunchKernelGGL(mlp_tobf16_k, dim3(nk_grid((int64_t)kHID * kKIN, 256)),
                     dim3(256), 0, stream(), (int64_t)kHID * kKIN,
                     s->th + kOffW1, s->w1b);
  NK_LAUNCHED();
  hipLaunchKernelGGL(mlp_tobf16_k, dim3(nk_grid((int64_t)kHID * kHID, 256)),
                     dim3(256), 0, stream(), (int64_t)kHID * kHID,
                     s->th + kOffW2, s->w2b);
  NK_LAUNCHED();
  hipLaunchKernelGGL(mlp_tobf16_k, dim3(nk_grid((int64_t)kNP * kHID, 256)),
                     dim3(256), 0, stream(), (int64_t)kNP * kHID,
                     s->th + kOffW3, s->w3b);
  NK_LAUNCHED();
  return kOk;
}

// q = mask(mult QQ^T N(r)): the network's element-local prediction for the
// residual r (its features scaled by the RMS of r, the output by it again),
// made continuous and zero on the mask -- the learned correction of the
// flexible CG (fcg.hip).  Uses the state's inference buffers; lx = 8.
int mlp_direction(void* mlp, int64_t n, const double* r, double* q,
                  const double* mult, void* gs, const int32_t* mask,
                  int64_t nmask) {
  MlpState* s = as_mlp(mlp);
  if (!s) return fail_arg("mlp_direction", "null network");
  if (s->n != n)
    return fail_arg("mlp_direction", "the network was made for another mesh");
  int rc = rms(s, r, s->sc);
  if (!rc) rc = make_features(s, r, s->sc, s->xb);
  if (!rc) rc = forward(s, s->xb, s->nelv, false, nullptr, q, s->sc);
  if (!rc) rc = gs_launch<double>(gs, q, NEKOAMD_GS_ADD, nullptr);
  if (rc) return rc;
  hipLaunchKernelGGL(mlp_mulmask_k, dim3(nk_grid(s->n, 256)), dim3(256), 0,
                     stream(), s->n, q, mult);
  NK_LAUNCHED();
  return bc_zero_launch<double>(q, mask, nmask, nullptr);
}

}  // namespace nekoamd

using namespace nekoamd;

extern "C" {

int nekoamd_mlp_create(void** mlp, const int* nx, const int* ny, const int* nz,
                       const int* lx, const int* capacity, const int* batch,
                       const int64_t* seed) {
  NK_INIT();
  NK_REQUIRE(mlp && nx && ny && nz && lx && capacity && batch && seed,
             "null argument");
  NK_REQUIRE(*lx == 8, "the network takes lx = 8 elements (512 GLL values)");
  NK_REQUIRE(*nx > 0 && *ny > 0 && *nz > 0, "empty mesh");
  NK_REQUIRE(*capacity >= 0 && *batch >= 32 && *batch % 32 == 0,
             "capacity >= 0 and batch a positive multiple of 32");
  MlpState* s = new MlpState();
  s->nx = *nx;
  s->ny = *ny;
  s->nz = *nz;
  s->nelv = *nx * *ny * *nz;
  s->n = (int64_t)s->nelv * kNP;
  s->cap = *capacity;
  s->rows = s->nelv > *batch ? s->nelv : *batch;
  const int64_t R = s->rows;
  s->ws_len = 32LL * kHID * kKIN;
  if ((int64_t)kNP * kHID * 32 > s->ws_len) s->ws_len = (int64_t)kNP * kHID * 32;
  hipError_t e = hipSuccess;
  s->nbr = dalloc<int32_t>((int64_t)s->nelv * 6, e);
  s->th = dalloc<float>(kNParPad, e);
  s->g = dalloc<float>(kNParPad, e);
  s->m = dalloc<float>(kNParPad, e);
  s->v = dalloc<float>(kNParPad, e);
  s->w1b = dalloc<uint16_t>((int64_t)kHID * kKIN, e);
  s->w2b = dalloc<uint16_t>((int64_t)kHID * kHID, e);
  s->w3b = dalloc<uint16_t>((int64_t)kNP * kHID, e);
  s->Xd = dalloc<uint16_t>((int64_t)s->cap * s->nelv * kKIN, e);
  s->Td = dalloc<float>((int64_t)s->cap * s->nelv * kNP, e);
  s->xb = dalloc<uint16_t>(R * kKIN, e);
  s->h1 = dalloc<uint16_t>(R * kHID, e);
  s->h2 = dalloc<uint16_t>(R * kHID, e);
  s->dzb = dalloc<uint16_t>(R * (kNP + kHID), e);
  s->z1 = dalloc<float>(R * kHID, e);
  s->z2 = dalloc<float>(R * kHID, e);
  s->y = dalloc<float>(R * kNP, e);
  s->dzf = dalloc<float>(R * (kNP + kHID), e);
  s->ws = dalloc<float>(s->ws_len, e);
  s->mean = dalloc<double>(s->nelv, e);
  s->sc = dalloc<double>(16, e);
  s->part = dalloc<double>(3 * kRedGrid, e);
  s->q = dalloc<double>(s->n, e);
  s->aq = dalloc<double>(s->n, e);
  s->bsave = dalloc<double>(s->n, e);
  if (e == hipSuccess)
    e = hipHostMalloc((void**)&s->host, 16 * sizeof(double),
                      hipHostMallocDefault);
  if (e != hipSuccess) {
    nekoamd_mlp_free(s);
    return fail_hip(__func__, e);
  }
  // neighbour table of the x-fastest box numbering
  std::vector<int32_t> nb((size_t)s->nelv * 6);
  for (int el = 0; el < s->nelv; ++el) {
    const int ex = el % s->nx, ey = (el / s->nx) % s->ny,
              ez = el / (s->nx * s->ny);
    const int d[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0},
                         {0, 1, 0},  {0, 0, -1}, {0, 0, 1}};
    for (int k = 0; k < 6; ++k) {
      const int jx = ex + d[k][0], jy = ey + d[k][1], jz = ez + d[k][2];
      const bool in = jx >= 0 && jx < s->nx && jy >= 0 && jy < s->ny &&
                      jz >= 0 && jz < s->nz;
      nb[(size_t)el * 6 + k] = in ? jx + s->nx * (jy + s->ny * jz) : -1;
    }
  }
  // initial weights U(-1/sqrt(fan_in), 1/sqrt(fan_in)), biases 0, padding 0
  std::vector<float> th((size_t)kNParPad, 0.0f);
  std::mt19937_64 rng((uint64_t)*seed);
  std::uniform_real_distribution<double> uni(-1.0, 1.0);
  for (int o = 0; o < kHID; ++o)
    for (int i = 0; i < kNIN; ++i)
      th[kOffW1 + (int64_t)o * kKIN + i] = (float)(uni(rng) / std::sqrt(518.0));
  for (int64_t i = 0; i < (int64_t)kHID * kHID; ++i)
    th[kOffW2 + i] = (float)(uni(rng) / 16.0);
  for (int64_t i = 0; i < (int64_t)kNP * kHID; ++i)
    th[kOffW3 + i] = (float)(uni(rng) / 16.0);
  e = hipMemcpy(s->nbr, nb.data(), nb.size() * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(s->th, th.data(), th.size() * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(s->m, 0, kNParPad * 4);
  if (e == hipSuccess) e = hipMemset(s->v, 0, kNParPad * 4);
  if (e == hipSuccess) e = hipMemset(s->g, 0, kNParPad * 4);
  if (e != hipSuccess) {
    nekoamd_mlp_free(s);
    return fail_hip(__func__, e);
  }
  int rc = refresh_bf16(s);
  if (rc) {
    nekoamd_mlp_free(s);
    return rc;
  }
  *mlp = s;
  return kOk;
}

int nekoamd_mlp_free(void* mlp) {
  MlpState* s = as_mlp(mlp);
  if (!s) return kOk;
  if (initialised()) NK_HIP(hipStreamSynchronize(stream()));
  void* dev[] = {s->nbr, s->th,  s->g,    s->m,    s->v,  s->w1b, s->w2b,
                 s->w3b, s->Xd,  s->Td,   s->xb,   s->h1, s->h2,  s->dzb,
                 s->z1,  s->z2,  s->y,    s->dzf,  s->ws, s->mean, s->sc,
                 s->part, s->q,  s->aq,   s->bsave};
  for (void* p : dev)
    if (p) NK_HIP(hipFree(p));
  if (s->host) NK_HIP(hipHostFree(s->host));
  delete s;
  return kOk;
}

int nekoamd_mlp_get_params(void* mlp, float* out) {
  NK_INIT();
  MlpState* s = as_mlp(mlp);
  NK_REQUIRE(s && out, "null argument");
  std::vector<float> th((size_t)kNParPad);
  NK_HIP(hipStreamSynchronize(stream()));
  NK_HIP(hipMemcpy(th.data(), s->th, kNParPad * 4, hipMemcpyDeviceToHost));
  int64_t o = 0;
  for (int r = 0; r < kHID; ++r)
    for (int i = 0; i < kNIN; ++i) out[o++] = th[kOffW1 + (int64_t)r * kKIN + i];
  for (int64_t i = kOffB1; i < kNParPad; ++i) out[o++] = th[i];
  return kOk;
}

int nekoamd_mlp_set_params(void* mlp, const float* in) {
  NK_INIT();
  MlpState* s = as_mlp(mlp);
  NK_REQUIRE(s && in, "null argument");
  std::vector<float> th((size_t)kNParPad, 0.0f);
  int64_t o = 0;
  for (int r = 0; r < kHID; ++r)
    for (int i = 0; i < kNIN; ++i) th[kOffW1 + (int64_t)r * kKIN + i] = in[o++];
  for (int64_t i = kOffB1; i < kNParPad; ++i) th[i] = in[o++];
  NK_HIP(hipStreamSynchronize(stream()));
  NK_HIP(hipMemcpy(s->th, th.data(), kNParPad * 4, hipMemcpyHostToDevice));
  NK_HIP(hipMemset(s->m, 0, kNParPad * 4));
  NK_HIP(hipMemset(s->v, 0, kNParPad * 4));
  s->adam_t = 0;
  return refresh_bf16(s);
}

int nekoamd_mlp_forward(void* mlp, const void* X, const int* M, void* Y,
                        void* H1, void* H2) {
  NK_INIT();
  MlpState* s = as_mlp(mlp);
  NK_REQUIRE(s && X && M && Y, "null argument");
  NK_REQUIRE(*M >= 0 && *M <= s->rows, "M exceeds the work buffers");
  int rc = forward(s, (const uint16_t*)X, *M, false, (float*)Y, nullptr,
                   nullptr);
  if (rc) return rc;
  if (H1)
    NK_HIP(hipMemcpyAsync(H1, s->h1, (size_t)*M * kHID * 2,
                          hipMemcpyDeviceToDevice, stream()));
  if (H2)
    NK_HIP(hipMemcpyAsync(H2, s->h2, (size_t)*M * kHID * 2,
                          hipMemcpyDeviceToDevice, stream()));
  return kOk;
}

int nekoamd_mlp_loss_grad(void* mlp, const void* X, const void* T,
                          const int* M, double* loss, void* g) {
  NK_INIT();
  MlpState* s = as_mlp(mlp);
  NK_REQUIRE(s && X && T && M && loss && g, "null argument");
  NK_REQUIRE(*M > 0 && *M <= s->rows && *M % 8 == 0,
             "M must be a positive multiple of 8 within the work buffers");
  // targets unscaled: sc[2..3] = {1, 1}
  s->host[0] = 1.0;
  s->host[1] = 1.0;
  NK_HIP(hipMemcpyAsync(s->sc + 2, s->host, 2 * sizeof(double),
                        hipMemcpyHostToDevice, stream()));
  NK_HIP(hipMemsetAsync(s->sc + 6, 0, sizeof(double), stream()));
  int np = 0;
  int rc = backward(s, (const uint16_t*)X, (const float*)T, *M, s->sc + 2,
                    &np);
  if (rc) return rc;
  hipLaunchKernelGGL(mlp_loss_acc_k, dim3(1), dim3(kFinBlock), 0, stream(),
                     s->part, np, s->sc + 6);
  NK_LAUNCHED();
  hipLaunchKernelGGL(mlp_unpad_k, dim3(nk_grid(NEKOAMD_MLP_NPARAM, 256)),
                     dim3(256), 0, stream(), s->g, (float*)g);
  NK_LAUNCHED();
  NK_HIP(hipMemcpyAsync(s->host + 2, s->sc + 6, sizeof(double),
                        hipMemcpyDeviceToHost, stream()));
  NK_HIP(hipStreamSynchronize(stream()));
  *loss = s->host[2] / ((double)*M * kNP);
  return kOk;
}

int nekoamd_mlp_harvest(void* mlp, const void* b, const void* x) {
  NK_INIT();
  MlpState* s = as_mlp(mlp);
  NK_REQUIRE(s && b && x, "null argument");
  if (s->cap == 0) return kOk;
  const int slot = s->next;
  uint16_t* X = s->Xd + (int64_t)slot * s->nelv * kKIN;
  float* T = s->Td + (int64_t)slot * s->n;
  int rc = rms(s, (const double*)b, s->sc);
  if (!rc) rc = make_features(s, (const double*)b, s->sc, X);
  if (rc) return rc;
  hipLaunchKernelGGL(mlp_target_k, dim3(nk_grid(s->n, 256)), dim3(256), 0,
                     stream(), s->n, (const double*)x, s->sc, T);
  NK_LAUNCHED();
  s->next = (s->next + 1) % s->cap;
  if (s->npairs < s->cap) ++s->npairs;
  return kOk;
}

int nekoamd_mlp_pairs(void* mlp, int* npairs) {
  MlpState* s = as_mlp(mlp);
  if (!s || !npairs) return fail_arg(__func__, "null argument");
  *npairs = s->npairs;
  return kOk;
}

int nekoamd_mlp_train(void* mlp, const int* epochs, const int* batch,
                      const float* lr, const int64_t* seed, double* loss_first,
                      double* loss_last) {
  NK_INIT();
  MlpState* s = as_mlp(mlp);
  NK_REQUIRE(s && epochs && batch && lr && seed && loss_first && loss_last,
             "null argument");
  NK_REQUIRE(*batch >= 32 && *batch % 32 == 0 && *batch <= s->rows,
             "batch must be a multiple of 32 within the work buffers");
  NK_REQUIRE(*epochs >= 1, "epochs must be positive");
  NK_REQUIRE(s->npairs > 0, "no harvested pairs");
  const int64_t S = (int64_t)s->npairs * s->nelv;
  const int B = *batch;
  const int64_t nb = S / B;
  NK_REQUIRE(nb > 0, "fewer samples than one batch");
  // target scale sc[2..3] = RMS of the stored targets and its inverse
  {
    const int64_t nt = S * kNP;
    const int g = nk_grid(nt, kRedBlock, kRedGrid);
    hipLaunchKernelGGL(mlp_sumsq_f_k, dim3(g), dim3(kRedBlock), 0, stream(),
                       nt, s->Td, s->part);
    NK_LAUNCHED();
    hipLaunchKernelGGL(mlp_rms_fin_k, dim3(1), dim3(kFinBlock), 0, stream(),
                       s->part, g, (double)nt, s->sc + 2);
    NK_LAUNCHED();
  }
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  std::mt19937_64 rng((uint64_t)*seed);
  std::vector<int64_t> order((size_t)nb);
  double* acc = s->sc + 6;
  for (int ep = 0; ep < *epochs; ++ep) {
    for (int64_t q = 0; q < nb; ++q) order[(size_t)q] = q;
    std::shuffle(order.begin(), order.end(), rng);
    NK_HIP(hipMemsetAsync(acc, 0, sizeof(double), stream()));
    for (int64_t q = 0; q < nb; ++q) {
      const int64_t r0 = order[(size_t)q] * B;
      int np = 0;
      int rc = backward(s, s->Xd + r0 * kKIN, s->Td + r0 * kNP, B, s->sc + 2,
                        &np);
      if (rc) return rc;
      hipLaunchKernelGGL(mlp_loss_acc_k, dim3(1), dim3(kFinBlock), 0,
                         stream(), s->part, np, acc);
      NK_LAUNCHED();
      ++s->adam_t;
      const float c1 = 1.0f - powf(b1, (float)s->adam_t);
      const float c2 = 1.0f - powf(b2, (float)s->adam_t);
      hipLaunchKernelGGL(mlp_adam_k, dim3(nk_grid(kNParPad, 256)), dim3(256),
                         0, stream(), s->th, s->g, s->m, s->v, *lr, b1, b2, c1,
                         c2, eps, s->w1b, s->w2b, s->w3b);
      NK_LAUNCHED();
    }
    NK_HIP(hipMemcpyAsync(s->host, acc, sizeof(double), hipMemcpyDeviceToHost,
                          stream()));
    NK_HIP(hipStreamSynchronize(stream()));
    const double l = s->host[0] / ((double)nb * B * kNP);
    if (ep == 0) *loss_first = l;
    *loss_last = l;
  }
  return kOk;
}

// Learned initial guess of one solve.  save: keep b for harvesting in
// nekoamd_mlp_post; apply: q = mask(mult QQ^T N(b)), aq = mask(QQ^T A q),
// gamma = <q,b>/<q,aq>, b -= gamma aq.
int nekoamd_mlp_pre(void* mlp, void* b, const void* mult, void* gs,
                    const void* mask, const int64_t* nmask, const void* dx,
                    const void* h1, const void* h2, const void* g11,
                    const void* g12, const void* g13, const void* g22,
                    const void* g23, const void* g33, const void* B,
                    const int* lx, const int* ifh2, const int* apply,
                    const int* save, double* gamma) {
  NK_INIT();
  MlpState* s = as_mlp(mlp);
  NK_REQUIRE(s && b && mult && gs && nmask && dx && h1 && g11 && g12 && g13 &&
                 g22 && g23 && g33 && lx && ifh2 && apply && save && gamma,
             "null argument");
  NK_REQUIRE(*lx == 8, "lx must be 8");
  NK_REQUIRE(*nmask >= 0 && (*nmask == 0 || mask), "bad mask");
  NK_REQUIRE(!*ifh2 || (h2 && B), "ifh2 set but h2 or B is null");
  double* bd = (double*)b;
  const int64_t n = s->n;
  const int ge = nk_grid(n, 256);
  *gamma = 0.0;
  s->applied = false;
  s->saved = false;
  if (*save) {
    NK_HIP(hipMemcpyAsync(s->bsave, bd, sizeof(double) * n,
                          hipMemcpyDeviceToDevice, stream()));
    s->saved = true;
  }
  if (!*apply) return kOk;
  int rc = rms(s, bd, s->sc);
  if (!rc) rc = make_features(s, bd, s->sc, s->xb);
  if (!rc) rc = forward(s, s->xb, s->nelv, false, nullptr, s->q, s->sc);
  if (!rc) rc = gs_launch<double>(gs, s->q, NEKOAMD_GS_ADD, nullptr);
  if (rc) return rc;
  hipLaunchKernelGGL(mlp_mulmask_k, dim3(ge), dim3(256), 0, stream(), n, s->q,
                     (const double*)mult);
  NK_LAUNCHED();
  rc = bc_zero_launch<double>(s->q, (const int32_t*)mask, *nmask, nullptr);
  if (!rc)
    rc = ax_helm_launch<double, double>(
        s->aq, s->q, (const double*)dx, (const double*)h1, (const double*)h2,
        (const double*)g11, (const double*)g12, (const double*)g13,
        (const double*)g22, (const double*)g23, (const double*)g33,
        (const double*)B, 8, s->nelv, *ifh2 != 0, nullptr);
  if (!rc) rc = gs_launch<double>(gs, s->aq, NEKOAMD_GS_ADD, nullptr);
  if (!rc)
    rc = bc_zero_launch<double>(s->aq, (const int32_t*)mask, *nmask, nullptr);
  if (rc) return rc;
  const int gr = nk_grid(n, kRedBlock, kRedGrid);
  hipLaunchKernelGGL(mlp_dots_k, dim3(gr), dim3(kRedBlock), 0, stream(), n,
                     s->q, bd, s->aq, (const double*)mult, s->part);
  NK_LAUNCHED();
  hipLaunchKernelGGL(mlp_gamma_k, dim3(1), dim3(kFinBlock), 0, stream(),
                     s->part, gr, s->sc + 8);
  NK_LAUNCHED();
  hipLaunchKernelGGL(mlp_axpy_k, dim3(ge), dim3(256), 0, stream(), n, bd,
                     s->aq, s->sc + 12);
  NK_LAUNCHED();
  NK_HIP(hipMemcpyAsync(s->host, s->sc + 8, 4 * sizeof(double),
                        hipMemcpyDeviceToHost, stream()));
  NK_HIP(hipStreamSynchronize(stream()));
  *gamma = s->host[3];
  s->applied = true;
  return kOk;
}

// x += gamma q when a guess was applied; harvest (kept b, x) when asked and
// b was kept.
int nekoamd_mlp_post(void* mlp, void* x, const int* harvest) {
  NK_INIT();
  MlpState* s = as_mlp(mlp);
  NK_REQUIRE(s && x && harvest, "null argument");
  if (s->applied) {
    hipLaunchKernelGGL(mlp_axpy_k, dim3(nk_grid(s->n, 256)), dim3(256), 0,
                       stream(), s->n, (double*)x, s->q, s->sc + 11);
    NK_LAUNCHED();
    s->applied = false;
  }
  if (*harvest && s->saved) {
    s->saved = false;
    return nekoamd_mlp_harvest(mlp, s->bsave, x);
  }
  return kOk;
}

}  // extern "C"
