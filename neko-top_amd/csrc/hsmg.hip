// hsmg.hip -- hybrid Schwarz multigrid preconditioner, Neko's pc_hsmg (the
// 'hsmg' of the shipped cases' pressure solver, rugby_ball.case:46-52; built
// by precon_factory next to 'jacobi', adjoint_scheme.f90:1077-1087).  Neko is
// not vendored; the method is that of Lottes & Fischer, J. Sci. Comput. 24
// (2005), with Neko's three levels:
//
//   fine lx, mid lx_mid = 4 (lx >= 5, else max(lx - 1, 3)), coarse lx = 2
//   z  = S_fine(r)
//   r2 = gs_mid(J_fm^T (mult r));                e2 = S_mid(r2)
//   r1 = mask(gs_crs(J_mc^T (mult_mid r2)));     e1 = mask(PCG_crs(r1)), 10 its
//   e2 += J_mc e1;  z += J_fm e2;  z = mult gs_fine(z)
// (on the device the last line is z = zf + J_fm e2 in one pass, zf the
// weighted fine Schwarz result: both terms are continuous, so the closing
// mult gs is the identity up to rounding)
//
// S is overlapping Schwarz with one-point overlap: every element extended
// by the neighbours' first interior layer, local solves by fast
// diagonalisation, weighted by the inverse overlap count.  The mesh is a
// box (element e = ex + nx (ey + ny ez)), so the ghost layers move by
// direct reads of the face neighbours' data: the extension reads the
// neighbours' first interior layers from r, the return adds the neighbours'
// local solutions at their ghost points onto this element's layers next to
// its faces.  (The same exchange as a gs on the (lx + 2)-point lattice whose
// ghost points coincide with the neighbours' opposite ghosts -- ghost :=
// own layer, gs, ghost -= own layer -- up to the rounding of (a + b) - a,
// without the two gs passes over 1.95 x the field.)
//
// Device work per Schwarz application: two element passes, the FDM on the
// extended element (built in LDS from the element and its neighbours'
// layers), and one that returns the neighbours' ghosts, sums the point's
// copies (the gs), applies the mask and the weight (the inverse overlap
// count, counted there from the box).  The FDM is six tensor passes of length lx + 2 over the extended
// element in LDS with the element's three eigenvector matrices, and a
// scale by 1 / (lam_r + lam_s + lam_t).
//
// Precision: the preconditioner's data path (Schwarz work arrays, FDM
// matrices and arithmetic, weights, transfers, the fine and mid gs) is FP64
// or FP32 (NEKOAMD_CG_MIXED), the coarse PCG FP64 either way.  In FP32 it is
// an inexact preconditioner of the FP64 flexible GMRES, which stops on its
// FP64 residual.
//
// Set-up on the host: per element and direction the 1-D extended operators
// (the element's GLL stiffness and lumped mass scaled by its length, plus
// each neighbour's matrices on its two points next to the shared face; the
// ghost point dropped at the domain boundary and the boundary point too on
// a Dirichlet face), the generalised eigenproblem by cyclic Jacobi on
// B^-1/2 A B^-1/2, lengths = the mean of the element's four edges per
// direction.
#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

#include <cstdlib>

#include "internal.h"
#include "nekoamd.h"
#include "reduce.h"

namespace nekoamd {
namespace {

constexpr int kHB = 256;

struct Level {
  int lx = 0, lxe = 0;
  int64_t n = 0, next = 0;
  void* gs = nullptr;
  int32_t* mask = nullptr;
  int64_t nmask = 0;
  void* S = nullptr;         // T [nelv][3][lxe][lxe], row = point, col = mode
  void* lam = nullptr;       // T [nelv][3][lxe], +inf for inactive modes
  void* W2 = nullptr;
};

struct Hsmg {
  int nx = 0, ny = 0, nz = 0, nelv = 0, lx = 0, lxm = 0, faces = 0;
  bool f32 = false;  // data path in FP32
  Level lv[3];  // 0 coarse, 1 mid, 2 fine
  double* J_fm = nullptr;  // [lx][lxm]: l^mid_j(z^fine_i)
  double* J_mc = nullptr;  // [lxm][2]
  // coarse level operator (FP64; the element arrays are set-up only)
  double *g[6] = {}, *B = nullptr, *dx = nullptr, *h1 = nullptr, *d = nullptr;
  int crs_iter = 10;
  // the coarse PCG on the assembled vertex lattice, (nx+1)(ny+1)(nz+1)
  // vertices a + (nx+1)(b + (ny+1)c): the 27 coefficients of each vertex's
  // row of the assembled operator ([27][nl], rows of masked vertices zero),
  // the inverse diagonal (zero on masked vertices), the CG vectors and the
  // partial sums of the 2 crs_iter reductions (crs_G per reduction)
  int64_t nl = 0;
  int crs_G = 0;
  double *S27 = nullptr, *dl = nullptr, *xl = nullptr, *rl = nullptr,
         *zl = nullptr, *wl = nullptr, *pl[2] = {}, *cpart = nullptr;
  // the coarse solve (fixed kernels on fixed buffers), recorded as a graph on
  // the first application and replayed (NEKOAMD_HIP_GRAPHS=1)
  hipGraphExec_t crs_graph = nullptr;
  // the graph runs on a side stream (default priority), overlapping the
  // fine and mid Schwarz steps: r1 ready -> ev_r1; coarse done -> ev_crs
  hipStream_t side = nullptr;
  hipEvent_t ev_r1 = nullptr, ev_crs = nullptr;
  bool graphs = false;  // NEKOAMD_HIP_GRAPHS=1 when the object was created
  double *r1 = nullptr, *e1 = nullptr;  // coarse, FP64
  void *r2 = nullptr, *e2 = nullptr;    // mid, T
  void *rf = nullptr, *zf = nullptr;    // fine, T
};

// ---- element kernels -------------------------------------------------------

// Whether local point (i, j, k) of the box element (ex, ey, ez) lies on one
// of the Dirichlet faces in `faces` (the points nekoamd_box_face_mask
// lists).
template <int LX>
__device__ __forceinline__ bool box_masked(int faces, int nx, int ny, int nz,
                                           int ex, int ey, int ez, int i,
                                           int j, int k) {
  return ((faces & NEKOAMD_FACE_XMIN) && ex == 0 && i == 0) ||
         ((faces & NEKOAMD_FACE_XMAX) && ex == nx - 1 && i == LX - 1) ||
         ((faces & NEKOAMD_FACE_YMIN) && ey == 0 && j == 0) ||
         ((faces & NEKOAMD_FACE_YMAX) && ey == ny - 1 && j == LX - 1) ||
         ((faces & NEKOAMD_FACE_ZMIN) && ez == 0 && k == 0) ||
         ((faces & NEKOAMD_FACE_ZMAX) && ez == nz - 1 && k == LX - 1);
}

// Point p of the extended element e of the masked r (zero on the Dirichlet
// faces `faces`): its own points, and in each face's ghost layer the face
// neighbour's first interior layer (lx - 2 below, 1 above), read straight
// from r (box numbering e = ex + nx (ey + ny ez)); zero on the domain
// boundary and on the ghost edges and corners.
template <int LX, typename T>
__device__ __forceinline__ T sch_ext_point(const T* __restrict__ r, int e,
                                           int ex, int ey, int ez, int nx,
                                           int ny, int nz, int faces, int p) {
  constexpr int LE = LX + 2, NP = LX * LX * LX;
  const int a = p % LE, b = (p / LE) % LE, c = p / (LE * LE);
  const bool ia = a >= 1 && a <= LX, ib = b >= 1 && b <= LX,
             ic = c >= 1 && c <= LX;
  // the element and local point read, if any
  int en = -1, qx = ex, qy = ey, qz = ez, i = a - 1, j = b - 1, k = c - 1;
  if (ia && ib && ic) {
    en = e;
  } else if (ib && ic && (a == 0 || a == LE - 1)) {
    const bool lo = a == 0;
    if (lo ? ex > 0 : ex < nx - 1) {
      en = lo ? e - 1 : e + 1;
      qx = lo ? ex - 1 : ex + 1;
      i = lo ? LX - 2 : 1;
    }
  } else if (ia && ic && (b == 0 || b == LE - 1)) {
    const bool lo = b == 0;
    if (lo ? ey > 0 : ey < ny - 1) {
      en = lo ? e - nx : e + nx;
      qy = lo ? ey - 1 : ey + 1;
      j = lo ? LX - 2 : 1;
    }
  } else if (ia && ib && (c == 0 || c == LE - 1)) {
    const bool lo = c == 0;
    if (lo ? ez > 0 : ez < nz - 1) {
      en = lo ? e - nx * ny : e + nx * ny;
      qz = lo ? ez - 1 : ez + 1;
      k = lo ? LX - 2 : 1;
    }
  }
  if (en < 0) return T(0);
  if (faces && box_masked<LX>(faces, nx, ny, nz, qx, qy, qz, i, j, k))
    return T(0);
  return r[(int64_t)en * NP + i + LX * (j + LX * k)];
}

// Threads of the FDM kernel for extended elements of LE points a side: one
// per line of a tensor pass (LE^2 lines), in whole wavefronts.
constexpr int fdm_threads(int le) { return (le * le + 63) / 64 * 64; }

// Outputs per (line, part) item of a tensor pass over lines of LE points.
constexpr int fdm_opt(int le) {
  return le % 5 == 0 ? 5 : le % 4 == 0 ? 4 : le % 3 == 0 ? 3 : le % 2 == 0 ? 2 : 1;
}

// LDS base and stride of line `line` along direction DIR (0: x, 1: y, 2: z)
// of an LE^3 block stored x fastest.
template <int LE, int DIR>
__device__ __forceinline__ void fdm_line(int line, int& base, int& str) {
  if (DIR == 0) {
    base = LE * line;
    str = 1;
  } else if (DIR == 1) {
    base = line % LE + LE * LE * (line / LE);
    str = LE;
  } else {
    base = line;
    str = LE * LE;
  }
}

// One tensor pass: out_line[i] = sum_l M(l, i) in_line[l] along DIR, M(l, i)
// = Mat[l LE + i] (TR false) or Mat[i LE + l] (TR true).  Each item holds
// its whole input line in registers and forms OPT outputs of it: 10 / OPT
// LDS reads of data per output instead of 10 (the matrix values are the
// same for the items of a wave that form the same outputs).  The sum runs
// over l in order.
template <int LE, int DIR, bool TR, typename T>
__device__ __forceinline__ void fdm_pass_lds(const T* in, T* out, const T* Mat) {
  constexpr int OPT = fdm_opt(LE), TPL = LE / OPT, NI = LE * LE * TPL;
  for (int item = threadIdx.x; item < NI; item += kHB) {
    const int line = item / TPL, part = item % TPL;
    int base, str;
    fdm_line<LE, DIR>(line, base, str);
    T x[LE];
#pragma unroll
    for (int l = 0; l < LE; ++l) x[l] = in[base + l * str];
#pragma unroll
    for (int o = 0; o < OPT; ++o) {
      const int i = part * OPT + o;
      T s = T(0);
#pragma unroll
      for (int l = 0; l < LE; ++l) s += Mat[TR ? i * LE + l : l * LE + i] * x[l];
      out[base + i * str] = s;
      // keep the matrix reads of the next output below this point: one
      // column of LE values live at a time
      asm volatile("" ::: "memory");
    }
  }
}

// p as a pointer into the constant address space, made opaque to the
// compiler by an empty asm that also takes `dep` (a value of the previous
// row's sums) and uniform by readfirstlane: loads through it are scalar
// loads (read-only memory, the same address in every lane) that cannot be
// issued before dep is computed, so one matrix row at a time is live in
// scalar registers.  p must be the same in every lane.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T*
uniform_row(const T* p, const T& dep) {
  uint64_t v = (uint64_t)p;
  asm volatile("" : "+s"(v) : "v"(dep));
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return (const __attribute__((address_space(4))) T*)(((uint64_t)hi << 32) |
                                                       lo);
}

// One tensor pass along DIR, one line of the LE^3 block per thread:
// out_line[i] = sum_l M(l, i) in_line[l], the sum over l in order, with
// M(l, i) = Me[l LE + i] (TR false) or Me[i LE + l] (TR true).  Me is the
// element's matrix in global memory: its index is the same for every thread
// of the workgroup, so the compiler reads it with scalar loads and feeds it
// to the FMAs as a scalar operand.  Per output one LDS read of data and one
// write (a matrix held in LDS costs a wave-wide LDS read per term).
template <int LE, int DIR, bool TR, typename T>
__device__ __forceinline__ void fdm_pass(const T* in, T* out,
                                         const T* __restrict__ Me) {
  const int line = threadIdx.x;
  if (line >= LE * LE) return;
  int base, str;
  fdm_line<LE, DIR>(line, base, str);
  T x[LE];
#pragma unroll
  for (int l = 0; l < LE; ++l) x[l] = in[base + l * str];
  // the matrix is walked row by row (contiguous), one row of LE values live
  // in scalar registers at a time (uniform_row keeps the compiler from
  // hoisting all LE^2 loads, which would not fit)
  if constexpr (!TR) {
    T y[LE];
#pragma unroll
    for (int i = 0; i < LE; ++i) y[i] = T(0);
#pragma unroll
    for (int l = 0; l < LE; ++l) {
      const auto row = uniform_row(Me + l * LE, y[LE - 1]);
#pragma unroll
      for (int i = 0; i < LE; ++i) y[i] += row[i] * x[l];
    }
#pragma unroll
    for (int i = 0; i < LE; ++i) out[base + i * str] = y[i];
  } else {
T prev = x[LE - 1];
#pragma unroll
    for (int i = 0; i < LE; ++i) {
      const auto row = uniform_row(Me + i * LE, prev);
      T s = T(0);
#pragma unroll
      for (int l = 0; l < LE; ++l) s += row[l] * x[l];
      out[base + i * str] = s;
      prev = s;
    }
  }
}

// The FDM solve of the extended element of r (built in LDS from the element
// and its face neighbours); W2 = the result.  Passes: x and
// y forward (S^T), then per z-line in registers the z forward pass, the
// scale by 1 / (lam_r + lam_s + lam_t) and the z backward pass (S), then y
// and x backward, the last one straight to W2.  Per output the same sums in
// the same order as six separate passes.  One thread per line, the
// element's matrices read through the scalar cache (uniform_row); the kernel
// for lx <= 6.
// Element of workgroup b in a launch with one workgroup per element (nb of
// them).  The dispatcher sends workgroup b to XCD b mod 8 (MI355X: 8 XCDs,
// each with its own L2); this mapping gives XCD x the contiguous run of
// elements [x per, (x + 1) per), so the face neighbours an element reads
// (its x and y neighbours, processed shortly before) are on the same XCD and
// their lines still in its L2.  A bijection for every nb.
__device__ __forceinline__ int xcd_element(int b, int nb) {
  const int per = nb >> 3;
  if (b >= per * 8) return b;
  return (b & 7) * per + (b >> 3);
}

template <int LX, typename T>
__global__ void __launch_bounds__(fdm_threads(LX + 2))
    sch_fdm_k(const T* __restrict__ r, T* __restrict__ W2, const T* __restrict__ S,
              const T* __restrict__ lam, int nx, int ny, int nz, int faces) {
  constexpr int LE = LX + 2, NE = LE * LE * LE, NT = fdm_threads(LE);
  __shared__ T sl[3][LE];
  __shared__ T u[NE], v[NE];
  const int e = xcd_element(blockIdx.x, gridDim.x);
  const int64_t eb = (int64_t)e * NE;
  const T* Se = S + (int64_t)e * 3 * LE * LE;  // S_x, S_y, S_z of the element
  for (int i = threadIdx.x; i < 3 * LE; i += NT)
    sl[i / LE][i % LE] = lam[(int64_t)e * 3 * LE + i];
  // warm L2 with the element's three matrices (one vector load per 64-byte
  // line, all in flight together while the extended element is built): the
  // passes then read them row by row through the scalar cache
  constexpr int kLine = 64 / (int)sizeof(T);
  static_assert(3 * LE * LE <= NT * kLine, "one line per thread");
  T warm = T(0);
  if ((int)threadIdx.x * kLine < 3 * LE * LE) warm = Se[threadIdx.x * kLine];
  {
    const int ex = e % nx, ey = (e / nx) % ny, ez = e / (nx * ny);
    // a compile-time trip count: every load of the thread is issued before
    // the first store (one memory latency, not one per point)
#pragma unroll
    for (int it = 0; it < (NE + NT - 1) / NT; ++it) {
      const int p = threadIdx.x + it * NT;
      if (p < NE)
        u[p] = sch_ext_point<LX, T>(r, e, ex, ey, ez, nx, ny, nz, faces, p);
    }
  }
  asm volatile("" ::"v"(warm));
  __syncthreads();
  fdm_pass<LE, 0, false>(u, v, Se);
  __syncthreads();
  fdm_pass<LE, 1, false>(v, u, Se + LE * LE);
  __syncthreads();
  // z forward, scale, z backward: one z-line per thread, in place in u
  const int ab = threadIdx.x;
  if (ab < LE * LE) {
    const T* Sz = Se + 2 * LE * LE;
    T x[LE], y[LE];
#pragma unroll
    for (int l = 0; l < LE; ++l) x[l] = u[ab + LE * LE * l];
    const T lab = sl[0][ab % LE] + sl[1][ab / LE];
    // rows of S_z one at a time, as in fdm_pass
#pragma unroll
    for (int c = 0; c < LE; ++c) y[c] = T(0);
#pragma unroll
    for (int l = 0; l < LE; ++l) {
      const auto row = uniform_row(Sz + l * LE, y[LE - 1]);
#pragma unroll
      for (int c = 0; c < LE; ++c) y[c] += row[c] * x[l];
      // all of the row's sums before the next row's loads (otherwise the
      // scheduler forms only y[LE - 1] early and loads every row at once)
#pragma unroll
      for (int c = 0; c < LE; ++c) asm volatile("" : "+v"(y[c]));
    }
#pragma unroll
    for (int c = 0; c < LE; ++c) {
      const T den = lab + sl[2][c];
      y[c] = (den != T(0) && isfinite(den)) ? y[c] / den : T(0);
    }
T prev = y[LE - 1];
#pragma unroll
    for (int c = 0; c < LE; ++c) {
      const auto row = uniform_row(Sz + c * LE, prev);
      T s = T(0);
#pragma unroll
      for (int l = 0; l < LE; ++l) s += row[l] * y[l];
      u[ab + LE * LE * c] = s;
      prev = s;
    }
  }
  __syncthreads();
  fdm_pass<LE, 1, true>(u, v, Se + LE * LE);
  __syncthreads();
  fdm_pass<LE, 0, true>(v, W2 + eb, Se);
}

// sch_fdm_k with the three matrices staged in LDS and kHB threads, each
// forming fdm_opt(LE) outputs of a line per item (the kernel for lx = 8).
template <int LX, typename T>
__global__ void __launch_bounds__(kHB)
    sch_fdm_lds_k(const T* __restrict__ r, T* __restrict__ W2, const T* __restrict__ S,
              const T* __restrict__ lam, int nx, int ny, int nz, int faces) {
  constexpr int LE = LX + 2, NE = LE * LE * LE;
  __shared__ T sS[3][LE * LE];
  __shared__ T sl[3][LE];
  __shared__ T u[NE], v[NE];
  const int e = xcd_element(blockIdx.x, gridDim.x);
  const int64_t eb = (int64_t)e * NE;
  for (int i = threadIdx.x; i < 3 * LE * LE; i += kHB)
    sS[i / (LE * LE)][i % (LE * LE)] = S[(int64_t)e * 3 * LE * LE + i];
  for (int i = threadIdx.x; i < 3 * LE; i += kHB)
    sl[i / LE][i % LE] = lam[(int64_t)e * 3 * LE + i];
  {
    const int ex = e % nx, ey = (e / nx) % ny, ez = e / (nx * ny);
    // a compile-time trip count: every load of the thread is issued before
    // the first store (one memory latency, not one per point)
#pragma unroll
    for (int it = 0; it < (NE + kHB - 1) / kHB; ++it) {
      const int p = threadIdx.x + it * kHB;
      if (p < NE)
        u[p] = sch_ext_point<LX, T>(r, e, ex, ey, ez, nx, ny, nz, faces, p);
    }
  }
  __syncthreads();
  fdm_pass_lds<LE, 0, false>(u, v, sS[0]);
  __syncthreads();
  fdm_pass_lds<LE, 1, false>(v, u, sS[1]);
  __syncthreads();
  // z forward, scale, z backward: one z-line per thread, in place in u
  for (int ab = threadIdx.x; ab < LE * LE; ab += kHB) {
    T x[LE], y[LE];
#pragma unroll
    for (int l = 0; l < LE; ++l) x[l] = u[ab + LE * LE * l];
    const T lab = sl[0][ab % LE] + sl[1][ab / LE];
#pragma unroll
    for (int c = 0; c < LE; ++c) {
      T s = T(0);
#pragma unroll
      for (int l = 0; l < LE; ++l) s += sS[2][l * LE + c] * x[l];
      const T den = lab + sl[2][c];
      y[c] = (den != T(0) && isfinite(den)) ? s / den : T(0);
      asm volatile("" ::: "memory");
    }
#pragma unroll
    for (int c = 0; c < LE; ++c) {
      T s = T(0);
#pragma unroll
      for (int l = 0; l < LE; ++l) s += sS[2][c * LE + l] * y[l];
      u[ab + LE * LE * c] = s;
      asm volatile("" ::: "memory");
    }
  }
  __syncthreads();
  fdm_pass_lds<LE, 1, true>(u, v, sS[1]);
  __syncthreads();
  fdm_pass_lds<LE, 0, true>(v, W2 + eb, sS[0]);
}

// One tensor pass of the FDM on the FP64 matrix cores: out_line(n)[i] =
// sum_l M(l, i) in_line(n)[l] for the LE^2 lines n along DIR, as the product
// D = A B with A[i][l] = M(l, i) (M(l, i) = Mat[l LE + i], or Mat[i LE + l]
// if TR) padded to 16 x 4KS and B[l][n] = in_line(n)[l] padded to 4KS x
// 16NT, in 16 x 16 tiles of v_mfma_f64_16x16x4f64 (lane t supplies
// A[t % 16][t / 16] and B[t / 16][t % 16] of each 4-deep step and holds
// D[4 i + t / 16][t % 16] in acc[i]).  The workgroup's waves take the column
// tiles in turn; padded rows and columns are zero and not written.
template <int LE, int DIR, bool TR>
__device__ __forceinline__ void fdm_pass_mfma(const double* in, double* out,
                                              const double* Mat) {
  typedef double f64x4 __attribute__((ext_vector_type(4)));
  constexpr int NL = LE * LE, NT = (NL + 15) / 16, KS = (LE + 3) / 4;
  constexpr int NW = kHB / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 15, q = lane >> 4;
  double a[KS];
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    const int k = 4 * ks + q;
    a[ks] = (r < LE && k < LE) ? (TR ? Mat[r * LE + k] : Mat[k * LE + r]) : 0.0;
  }
#pragma unroll
  for (int j = 0; j < (NT + NW - 1) / NW; ++j) {
    const int nt = wave + j * NW;  // the same in every lane of the wave
    if (nt >= NT) break;
    const int n = 16 * nt + r;
    int base = 0, str = 0;
    if (n < NL) fdm_line<LE, DIR>(n, base, str);
    f64x4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int k = 4 * ks + q;
      const double b = (n < NL && k < LE) ? in[base + k * str] : 0.0;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ks], b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int m = 4 * i + q;
      if (m < LE && n < NL) out[base + m * str] = acc[i];
    }
  }
}

// sch_fdm_k in FP64 with the six tensor passes on the matrix cores (the
// kernel for lx = 8 in FP64): the extended element in LDS, x, y, z forward,
// the scale by 1 / (lam_r + lam_s + lam_t) in place, z, y, x backward, the
// last pass straight to W2.  21 MFMAs of 16 x 16 x 4 per pass (the 10 x 10
// operator padded to 16 x 12, the 100 lines to 112); the LDS reads per pass
// are the 1000 inputs once, instead of about 12 wave-wide reads per output
// of the LDS kernel.  Each output sums in the matrix cores' order.
template <int LX>
__global__ void __launch_bounds__(kHB)
    sch_fdm_mfma_k(const double* __restrict__ r, double* __restrict__ W2,
                   const double* __restrict__ S, const double* __restrict__ lam,
                   int nx, int ny, int nz, int faces) {
  constexpr int LE = LX + 2, NE = LE * LE * LE;
  __shared__ double sl[3][LE];
  __shared__ double sS[3 * LE * LE];
  __shared__ double u[NE], v[NE];
  const int e = blockIdx.x;
  const int64_t eb = (int64_t)e * NE;
  // the three matrices staged in LDS with the extended element (one wait)
  for (int i = threadIdx.x; i < 3 * LE * LE; i += kHB)
    sS[i] = S[(int64_t)e * 3 * LE * LE + i];
  for (int i = threadIdx.x; i < 3 * LE; i += kHB)
    sl[i / LE][i % LE] = lam[(int64_t)e * 3 * LE + i];
  {
    const int ex = e % nx, ey = (e / nx) % ny, ez = e / (nx * ny);
    // a compile-time trip count: every load of the thread is issued before
    // the first store (one memory latency, not one per point)
#pragma unroll
    for (int it = 0; it < (NE + kHB - 1) / kHB; ++it) {
      const int p = threadIdx.x + it * kHB;
      if (p < NE)
        u[p] = sch_ext_point<LX, double>(r, e, ex, ey, ez, nx, ny, nz, faces, p);
    }
  }
  __syncthreads();
  fdm_pass_mfma<LE, 0, false>(u, v, sS);
  __syncthreads();
  fdm_pass_mfma<LE, 1, false>(v, u, sS + LE * LE);
  __syncthreads();
  fdm_pass_mfma<LE, 2, false>(u, v, sS + 2 * LE * LE);
  __syncthreads();
  for (int p = threadIdx.x; p < NE; p += kHB) {
    const int a = p % LE, b = (p / LE) % LE, c = p / (LE * LE);
    const double den = sl[0][a] + sl[1][b] + sl[2][c];
    v[p] = (den != 0.0 && isfinite(den)) ? v[p] / den : 0.0;
  }
  __syncthreads();
  fdm_pass_mfma<LE, 2, true>(v, u, sS + 2 * LE * LE);
  __syncthreads();
  fdm_pass_mfma<LE, 1, true>(u, v, sS + LE * LE);
  __syncthreads();
  fdm_pass_mfma<LE, 0, true>(v, W2 + eb, sS);
}

// The local-solution return, the gs over the point's copies, the mask and
// the weights in one pass.  For the global lattice position P of each local
// point: the sum over every element g whose extended domain holds P -- as
// its own point, or in one face's ghost layer (a ghost edge or corner
// never counts) -- of W2_g at P.  That is the per-element return (its own
// value plus the face neighbours' ghost values there) summed over the
// copies of P, as gs would; the terms are taken in one fixed order (z, then
// y, then x candidates ascending), so every copy of P gets the same value.
// Zero on the Dirichlet faces `faces`; with WT times the weight 1 / (the
// number of terms), the inverse overlap count of the method (the count of
// the same data movement applied to ones: every term's ghost source lies in
// the domain).
// Per dimension a lattice coordinate X lies in the extended range
// [L g - 1, L g + L + 1] (L = lx - 1) of at most C elements g, C = 3 at
// lx = 3 (an element's one interior layer is both of its neighbours' ghost
// layers) and 2 above; candidate c is g = (X + 1) / L - (C - 1) + c, valid
// when it exists and its range holds X, a ghost when X is its index 0 or
// lx + 1.  All C^3 terms are formed unrolled, the valid ones loaded (under
// the lane mask) before any is added, so a thread's loads are in flight
// together.
template <int LX>
struct SchCover {
  static constexpr int C = LX == 3 ? 3 : 2;
  int g[C], a[C];
  bool ok[C], gh[C];
  __device__ __forceinline__ SchCover(int X, int n) {
    constexpr int L = LX - 1, LE = LX + 2;
    const int g1 = (X + 1) / L;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int q = g1 - (C - 1) + c;
      const int ai = X - L * q + 1;
      ok[c] = q >= 0 && q < n && ai >= 0 && ai <= LE - 1;
      g[c] = q;
      a[c] = ai;
      gh[c] = ai == 0 || ai == LE - 1;
    }
  }
};

// One thread per point; workgroups of kHB threads hold kHB / lx^3 whole
// elements when an element has fewer points (the mid level, lx = 4: four
// elements per workgroup), one element of lx^3 threads (rounded to 64)
// when it has more (lx = 8: 131 us with 512 threads against 138 us with 256
// threads taking two points each).
template <int LX>
constexpr int sch_ret_gs_threads() {
  return LX * LX * LX > kHB ? (LX * LX * LX + 63) / 64 * 64 : kHB;
}
template <int LX>
constexpr int sch_ret_gs_epb() {
  return LX * LX * LX > kHB ? 1 : kHB / (LX * LX * LX);
}

template <int LX, typename T, bool WT>
__global__ void __launch_bounds__(sch_ret_gs_threads<LX>())
    sch_ret_gs_k(T* __restrict__ out, const T* __restrict__ W2, int nx, int ny,
                 int nz, int faces) {
  constexpr int LE = LX + 2, NE = LE * LE * LE, NP = LX * LX * LX, L = LX - 1;
  constexpr int C = SchCover<LX>::C, EPB = sch_ret_gs_epb<LX>();
  const int e = xcd_element(blockIdx.x, gridDim.x) * EPB + (int)threadIdx.x / NP;
  const int p = (int)threadIdx.x % NP;
  if ((int)threadIdx.x >= EPB * NP || e >= nx * ny * nz) return;
  const int ex = e % nx, ey = (e / nx) % ny, ez = e / (nx * ny);
  const int i = p % LX, j = (p / LX) % LX, k = p / (LX * LX);
  const int64_t o = (int64_t)e * NP + p;
  const SchCover<LX> cx(L * ex + i, nx), cy(L * ey + j, ny), cz(L * ez + k, nz);
  T t[C * C * C];
  int nt = 0;
#pragma unroll
  for (int qz = 0; qz < C; ++qz)
#pragma unroll
    for (int qy = 0; qy < C; ++qy)
#pragma unroll
      for (int qx = 0; qx < C; ++qx) {
        const bool use = cx.ok[qx] && cy.ok[qy] && cz.ok[qz] &&
                         (int)cx.gh[qx] + (int)cy.gh[qy] + (int)cz.gh[qz] <= 1;
        const int64_t g = cx.g[qx] + (int64_t)nx * (cy.g[qy] + (int64_t)ny * cz.g[qz]);
        const int q = qx + C * (qy + C * qz);
        t[q] = T(0);
        if (use) t[q] = W2[g * NE + cx.a[qx] + LE * (cy.a[qy] + LE * cz.a[qz])];
        nt += use ? 1 : 0;
      }
  T v = T(0);
#pragma unroll
  for (int q = 0; q < C * C * C; ++q) v += t[q];
  if (box_masked<LX>(faces, nx, ny, nz, ex, ey, ez, i, j, k)) v = T(0);
  if constexpr (WT) v = v * (T)(1.0 / (double)nt);
  out[o] = v;
}

// out = (TO) a
template <typename TO, typename TA>
__global__ void __launch_bounds__(kHB)
    hs_cast_k(int64_t n, TO* __restrict__ out, const TA* __restrict__ a) {
  for (int64_t p = (int64_t)blockIdx.x * kHB + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * kHB)
    out[p] = (TO)a[p];
}

// Elements per workgroup of the transfer kernels: kHB / (fine points) whole
// elements when an element has fewer fine points than kHB (the mid-to-coarse
// pair, lx 4 -> 2: four elements of 64 threads), else one.
template <int NF>
constexpr int hs_epb() {
  return NF < kHB ? kHB / NF : 1;
}

// out (coarse) = (J^T x J^T x J^T)(mult ⊙ r) per element, J [LF][LC].
// mult = 1 / (the point's number of copies) is counted from the box: the
// copies multiply by 2 per direction in which the point lies on a face
// shared with a neighbour (exactly the gs-built multiplicity, 1, 1/2, 1/4
// or 1/8, without reading it).
template <int LF, int LC, typename TI, typename TO>
__global__ void __launch_bounds__(kHB)
    hs_restrict_k(TO* __restrict__ out, const TI* __restrict__ r, int nx,
                  int ny, int nz, const double* __restrict__ J) {
  constexpr int NF = LF * LF * LF, EPB = hs_epb<NF>(), TPE = kHB / EPB;
  __shared__ double sJ[LF * LC];
  __shared__ double a[EPB][NF], b[EPB][LC * LF * LF], c[EPB][LC * LC * LF];
  const int el = threadIdx.x / TPE, lt = threadIdx.x % TPE;
  const int e = blockIdx.x * EPB + el;
  const bool live = el < EPB && e < nx * ny * nz;
  const int ex = e % nx, ey = (e / nx) % ny, ez = e / (nx * ny);
  for (int i = threadIdx.x; i < LF * LC; i += kHB) sJ[i] = J[i];
  // a compile-time trip count: all loads of the thread in flight together
  if (live) {
#pragma unroll
    for (int it = 0; it < (NF + TPE - 1) / TPE; ++it) {
      const int p = lt + it * TPE;
      if (p < NF) {
        const int i = p % LF, j = (p / LF) % LF, k = p / (LF * LF);
        const int m = ((i == 0 && ex > 0) || (i == LF - 1 && ex < nx - 1) ? 2 : 1) *
                      ((j == 0 && ey > 0) || (j == LF - 1 && ey < ny - 1) ? 2 : 1) *
                      ((k == 0 && ez > 0) || (k == LF - 1 && ez < nz - 1) ? 2 : 1);
        a[el][p] = (double)r[(int64_t)e * NF + p] * (1.0 / (double)m);
      }
    }
  }
  __syncthreads();
  if (live)
    for (int p = lt; p < LC * LF * LF; p += TPE) {
      const int q = p % LC, jk = p / LC;
      double s = 0.0;
      for (int l = 0; l < LF; ++l) s += sJ[l * LC + q] * a[el][l + LF * jk];
      b[el][p] = s;
    }
  __syncthreads();
  if (live)
    for (int p = lt; p < LC * LC * LF; p += TPE) {
      const int q = p % LC, r2 = (p / LC) % LC, k = p / (LC * LC);
      double s = 0.0;
      for (int l = 0; l < LF; ++l) s += sJ[l * LC + r2] * b[el][q + LC * (l + LF * k)];
      c[el][p] = s;
    }
  __syncthreads();
  if (live)
    for (int p = lt; p < LC * LC * LC; p += TPE) {
      const int qr = p % (LC * LC), t = p / (LC * LC);
      double s = 0.0;
      for (int l = 0; l < LF; ++l) s += sJ[l * LC + t] * c[el][qr + LC * LC * l];
      out[(int64_t)e * LC * LC * LC + p] = (TO)s;
    }
}

// out (fine) = w base + (J x J x J) ec per element (base when w is null),
// elements 0 .. nelv - 1
template <int LC, int LF, typename TI, typename TO, typename TB>
__global__ void __launch_bounds__(kHB)
    hs_prolong_k(TO* out, const TB* base, const TB* __restrict__ w,
                 const TI* __restrict__ ec, const double* __restrict__ J,
                 int nelv) {
  constexpr int NC = LC * LC * LC, NF = LF * LF * LF;
  constexpr int EPB = hs_epb<NF>(), TPE = kHB / EPB;
  constexpr int NB = (NF + TPE - 1) / TPE;
  __shared__ double sJ[LF * LC];
  __shared__ double a[EPB][NC], b[EPB][LF * LC * LC], c[EPB][LF * LF * LC];
  const int el = threadIdx.x / TPE, lt = threadIdx.x % TPE;
  const int e = blockIdx.x * EPB + el;
  const bool live = el < EPB && e < nelv;
  // base (x w) at this thread's fine points, loaded before the contractions
  // so that their latency is hidden (each point is read and written by the
  // same thread: safe when out is base)
  double bgv[NB];
#pragma unroll
  for (int it = 0; it < NB; ++it) {
    const int p = lt + it * TPE;
    bgv[it] = 0.0;
    if (live && p < NF) {
      const int64_t g = (int64_t)e * NF + p;
      bgv[it] = w ? (double)(base[g] * w[g]) : (double)base[g];
    }
  }
  for (int i = threadIdx.x; i < LF * LC; i += kHB) sJ[i] = J[i];
  if (live)
    for (int p = lt; p < NC; p += TPE) a[el][p] = (double)ec[(int64_t)e * NC + p];
  __syncthreads();
  if (live)
    for (int p = lt; p < LF * LC * LC; p += TPE) {
      const int q = p % LF, jk = p / LF;
      double s = 0.0;
      for (int l = 0; l < LC; ++l) s += sJ[q * LC + l] * a[el][l + LC * jk];
      b[el][p] = s;
    }
  __syncthreads();
  if (live)
    for (int p = lt; p < LF * LF * LC; p += TPE) {
      const int q = p % LF, r2 = (p / LF) % LF, k = p / (LF * LF);
      double s = 0.0;
      for (int l = 0; l < LC; ++l) s += sJ[r2 * LC + l] * b[el][q + LF * (l + LC * k)];
      c[el][p] = s;
    }
  __syncthreads();
  if (!live) return;
#pragma unroll
  for (int it = 0; it < NB; ++it) {
    const int p = lt + it * TPE;
    if (p >= NF) break;
    const int qr = p % (LF * LF), t = p / (LF * LF);
    double s = 0.0;
    for (int l = 0; l < LC; ++l) s += sJ[t * LC + l] * c[el][qr + LF * LF * l];
    out[(int64_t)e * NF + p] = (TO)(bgv[it] + s);
  }
}

// ---- host helpers ------------------------------------------------------------

#define HS_LX_SWITCH(LXV, CALL)                              \
  switch (LXV) {                                             \
    case 3: { constexpr int LV = 3; CALL; } break;           \
    case 4: { constexpr int LV = 4; CALL; } break;           \
    case 6: { constexpr int LV = 6; CALL; } break;           \
    case 8: { constexpr int LV = 8; CALL; } break;           \
    default: return fail_arg("hsmg", "unsupported level lx"); \
  }

template <typename TI, typename TO>
int restrict_launch(const Hsmg* h, int lf, int lc, TO* out, const TI* r,
                    const double* J) {
  const int ne = h->nelv, nx = h->nx, ny = h->ny, nz = h->nz;
  const dim3 b(kHB);
  auto go = [&](auto kern, int epb) {
    hipLaunchKernelGGL(kern, dim3((ne + epb - 1) / epb), b, 0, stream(), out, r,
                       nx, ny, nz, J);
  };
  if (lf == 8 && lc == 4)
    go(hs_restrict_k<8, 4, TI, TO>, hs_epb<512>());
  else if (lf == 6 && lc == 4)
    go(hs_restrict_k<6, 4, TI, TO>, hs_epb<216>());
  else if (lf == 4 && lc == 3)
    go(hs_restrict_k<4, 3, TI, TO>, hs_epb<64>());
  else if (lf == 4 && lc == 2)
    go(hs_restrict_k<4, 2, TI, TO>, hs_epb<64>());
  else if (lf == 3 && lc == 2)
    go(hs_restrict_k<3, 2, TI, TO>, hs_epb<27>());
  else
    return fail_arg("hsmg", "unsupported level pair");
  NK_LAUNCHED();
  return kOk;
}

// out = w base + J ec (w may be null: out = base + J ec; out may alias base)
template <typename TI, typename TO, typename TB>
int prolong_launch(int lc, int lf, int nelv, TO* out, const TB* base,
                   const TB* w, const TI* ec, const double* J) {
  const dim3 b(kHB);
  auto go = [&](auto kern, int epb) {
    hipLaunchKernelGGL(kern, dim3((nelv + epb - 1) / epb), b, 0, stream(), out,
                       base, w, ec, J, nelv);
  };
  if (lf == 8 && lc == 4)
    go(hs_prolong_k<4, 8, TI, TO, TB>, hs_epb<512>());
  else if (lf == 6 && lc == 4)
    go(hs_prolong_k<4, 6, TI, TO, TB>, hs_epb<216>());
  else if (lf == 4 && lc == 3)
    go(hs_prolong_k<3, 4, TI, TO, TB>, hs_epb<64>());
  else if (lf == 4 && lc == 2)
    go(hs_prolong_k<2, 4, TI, TO, TB>, hs_epb<64>());
  else if (lf == 3 && lc == 2)
    go(hs_prolong_k<2, 3, TI, TO, TB>, hs_epb<27>());
  else
    return fail_arg("hsmg", "unsupported level pair");
  NK_LAUNCHED();
  return kOk;
}

// One Schwarz application on level L: the FDM on the masked extended
// element, then the return, gs, mask and (weights) the weights in one pass
// (sch_ret_gs_k).  out may alias r.
template <typename T>
int schwarz(Hsmg* h, Level& L, T* out, const T* r, bool weights) {
  const int ne = h->nelv;
  T* W2 = (T*)L.W2;
  // the FDM reads r masked on the fly: no masked copy of r
  // lx <= 6: matrices through the scalar cache, one line per thread; lx = 8:
  // matrices in LDS (the scalar kernel's row loads are latency-bound there:
  // 415 against 316 us at 32^3, while at lx = 4 it takes 63 against 88 us)
  HS_LX_SWITCH(L.lx, if constexpr (LV <= 6) {
    hipLaunchKernelGGL((sch_fdm_k<LV, T>), dim3(ne), dim3(fdm_threads(LV + 2)),
                       0, stream(), r, W2, (const T*)L.S, (const T*)L.lam,
                       h->nx, h->ny, h->nz, h->faces);
  } else if constexpr (sizeof(T) == 8) {
    hipLaunchKernelGGL((sch_fdm_mfma_k<LV>), dim3(ne), dim3(kHB), 0, stream(),
                       (const double*)r, (double*)W2, (const double*)L.S,
                       (const double*)L.lam, h->nx, h->ny, h->nz, h->faces);
  } else {
    hipLaunchKernelGGL((sch_fdm_lds_k<LV, T>), dim3(ne), dim3(kHB), 0,
                       stream(), r, W2, (const T*)L.S, (const T*)L.lam, h->nx,
                       h->ny, h->nz, h->faces);
  });
  NK_LAUNCHED();
  if (weights) {
    HS_LX_SWITCH(L.lx, hipLaunchKernelGGL((sch_ret_gs_k<LV, T, true>),
                                           dim3((ne + sch_ret_gs_epb<LV>() - 1) /
                                                sch_ret_gs_epb<LV>()),
                                           dim3(sch_ret_gs_threads<LV>()), 0, stream(),
                                           out, W2, h->nx, h->ny, h->nz, h->faces));
  } else {
    HS_LX_SWITCH(L.lx, hipLaunchKernelGGL((sch_ret_gs_k<LV, T, false>),
                                           dim3((ne + sch_ret_gs_epb<LV>() - 1) /
                                                sch_ret_gs_epb<LV>()),
                                           dim3(sch_ret_gs_threads<LV>()), 0, stream(),
                                           out, W2, h->nx, h->ny, h->nz, h->faces));
  }
  NK_LAUNCHED();
  return kOk;
}

// host FP64 data -> a new device array of T
int upload_as(bool f32, void** dst, const std::vector<double>& src) {
  if (f32) {
    std::vector<float> f(src.begin(), src.end());
    NK_HIP(hipMalloc(dst, sizeof(float) * f.size()));
    NK_HIP(hipMemcpy(*dst, f.data(), sizeof(float) * f.size(), hipMemcpyHostToDevice));
  } else {
    NK_HIP(hipMalloc(dst, sizeof(double) * src.size()));
    NK_HIP(hipMemcpy(*dst, src.data(), sizeof(double) * src.size(),
                     hipMemcpyHostToDevice));
  }
  return kOk;
}

// Symmetric eigen-decomposition by cyclic Jacobi rotations: A (n x n, row
// major) is destroyed; lam[i], Q columns = eigenvectors.
void sym_eig(int n, std::vector<double>& A, std::vector<double>& lam,
             std::vector<double>& Q) {
  Q.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) Q[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        const double a = A[(size_t)i * n + j] * A[(size_t)i * n + j];
        tot += a;
        if (i != j) off += a;
      }
    if (off <= 1e-32 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = A[(size_t)p * n + q];
        if (apq == 0.0) continue;
        const double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
        const double th = (aqq - app) / (2.0 * apq);
        const double t = (th >= 0 ? 1.0 : -1.0) /
                         (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {  // columns p, q
          const double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
          A[(size_t)k * n + p] = c * akp - s * akq;
          A[(size_t)k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  // rows p, q
          const double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
          A[(size_t)p * n + k] = c * apk - s * aqk;
          A[(size_t)q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double qkp = Q[(size_t)k * n + p], qkq = Q[(size_t)k * n + q];
          Q[(size_t)k * n + p] = c * qkp - s * qkq;
          Q[(size_t)k * n + q] = s * qkp + c * qkq;
        }
      }
  }
  lam.resize(n);
  for (int i = 0; i < n; ++i) lam[i] = A[(size_t)i * n + i];
}

// 1-D GLL stiffness A_hat = D^T W D and weights of lx points
void one_d(int lx, std::vector<double>& w, std::vector<double>& Ah,
           std::vector<double>& z) {
  std::vector<double> D((size_t)lx * lx);
  z.resize(lx);
  w.resize(lx);
  nekoamd_gll(&lx, z.data(), w.data(), D.data());
  Ah.assign((size_t)lx * lx, 0.0);
  for (int i = 0; i < lx; ++i)
    for (int j = 0; j < lx; ++j) {
      double s = 0.0;
      for (int q = 0; q < lx; ++q)
        s += D[(size_t)q * lx + i] * w[q] * D[(size_t)q * lx + j];
      Ah[(size_t)i * lx + j] = s;
    }
}

// FDM data of one Schwarz level (host), uploaded to L.S and L.lam
int fdm_setup(Hsmg* h, Level& L, const std::vector<double>& len) {
  const int lx = L.lx, le = L.lxe, ne = h->nelv;
  std::vector<double> w, Ah, z;
  one_d(lx, w, Ah, z);
  std::vector<double> S((size_t)ne * 3 * le * le, 0.0),
      lam((size_t)ne * 3 * le, INFINITY);
  const int dims[3] = {h->nx, h->ny, h->nz};
  const int strides[3] = {1, h->nx, h->nx * h->ny};
  // elements are independent: split them over the host threads
  auto work = [&](int e0, int e1) {
  std::vector<double> A, Bm, C, lv, Q;
  for (int e = e0; e < e1; ++e) {
    const int ec[3] = {e % h->nx, (e / h->nx) % h->ny, e / (h->nx * h->ny)};
    for (int d = 0; d < 3; ++d) {
      const bool lo = ec[d] > 0, hi = ec[d] < dims[d] - 1;
      A.assign((size_t)le * le, 0.0);
      Bm.assign(le, 0.0);
      const double Le = len[(size_t)e * 3 + d];
      for (int i = 0; i < lx; ++i) {
        Bm[i + 1] += 0.5 * Le * w[i];
        for (int j = 0; j < lx; ++j)
          A[(size_t)(i + 1) * le + j + 1] += 2.0 / Le * Ah[(size_t)i * lx + j];
      }
      if (lo) {
        const double Ll = len[(size_t)(e - strides[d]) * 3 + d];
        for (int i = 0; i < 2; ++i) {
          Bm[i] += 0.5 * Ll * w[lx - 2 + i];
          for (int j = 0; j < 2; ++j)
            A[(size_t)i * le + j] += 2.0 / Ll * Ah[(size_t)(lx - 2 + i) * lx + lx - 2 + j];
        }
      }
      if (hi) {
        const double Lr = len[(size_t)(e + strides[d]) * 3 + d];
        for (int i = 0; i < 2; ++i) {
          Bm[le - 2 + i] += 0.5 * Lr * w[i];
          for (int j = 0; j < 2; ++j)
            A[(size_t)(le - 2 + i) * le + le - 2 + j] += 2.0 / Lr * Ah[(size_t)i * lx + j];
        }
      }
      std::vector<int> act;
      for (int i = 0; i < le; ++i) {
        bool on = true;
        if (!lo && (i == 0 || (i == 1 && (h->faces & (1 << (2 * d)))))) on = false;
        if (!hi && (i == le - 1 || (i == le - 2 && (h->faces & (1 << (2 * d + 1))))))
          on = false;
        if (on) act.push_back(i);
      }
      const int na = (int)act.size();
      C.assign((size_t)na * na, 0.0);
      for (int i = 0; i < na; ++i)
        for (int j = 0; j < na; ++j)
          C[(size_t)i * na + j] = A[(size_t)act[i] * le + act[j]] /
                                  std::sqrt(Bm[act[i]] * Bm[act[j]]);
      sym_eig(na, C, lv, Q);
      double lmax = 0.0;
      for (double x : lv) lmax = std::max(lmax, std::fabs(x));
      double* Sd = S.data() + ((size_t)e * 3 + d) * le * le;
      double* ld = lam.data() + ((size_t)e * 3 + d) * le;
      for (int k = 0; k < na; ++k) {
        ld[k] = std::fabs(lv[k]) < 1e-10 * lmax ? 0.0 : lv[k];
        for (int i = 0; i < na; ++i)
          Sd[(size_t)act[i] * le + k] = Q[(size_t)i * na + k] / std::sqrt(Bm[act[i]]);
      }
    }
  }
  };
  const int nt = std::max(1, std::min<int>(std::thread::hardware_concurrency(), 16));
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; ++t)
    pool.emplace_back(work, (int)((int64_t)ne * t / nt), (int)((int64_t)ne * (t + 1) / nt));
  for (auto& th : pool) th.join();
  int rc = upload_as(h->f32, &L.S, S);
  if (rc == kOk) rc = upload_as(h->f32, &L.lam, lam);
  return rc;
}

int level_setup(Hsmg* h, Level& L, int lx, bool schwarz_level,
                const std::vector<double>& len) {
  L.lx = lx;
  L.lxe = lx + 2;
  L.n = (int64_t)h->nelv * lx * lx * lx;
  std::vector<int64_t> gid(L.n);
  int rc = nekoamd_box_gids(&h->nx, &h->ny, &h->nz, &lx, gid.data());
  if (rc != kOk) return rc;
  if ((rc = nekoamd_gs_init(&L.gs, gid.data(), &L.n)) != kOk) return rc;
  int64_t cnt = 0;
  if ((rc = nekoamd_box_face_mask(&h->nx, &h->ny, &h->nz, &lx, &h->faces,
                                  nullptr, &cnt)) != kOk)
    return rc;
  L.nmask = cnt;
  std::vector<int32_t> m(std::max<int64_t>(cnt, 1));
  if ((rc = nekoamd_box_face_mask(&h->nx, &h->ny, &h->nz, &lx, &h->faces,
                                  m.data(), &cnt)) != kOk)
    return rc;
  NK_HIP(hipMalloc(&L.mask, sizeof(int32_t) * m.size()));
  NK_HIP(hipMemcpy(L.mask, m.data(), sizeof(int32_t) * m.size(),
                   hipMemcpyHostToDevice));
  const size_t es = h->f32 ? sizeof(float) : sizeof(double);
  if (!schwarz_level) return kOk;
  const int le = L.lxe;
  L.next = (int64_t)h->nelv * le * le * le;
  NK_HIP(hipMalloc(&L.W2, es * L.next));
  return fdm_setup(h, L, len);
}

void level_free(Level& L) {
  if (L.gs) nekoamd_gs_free(L.gs);
  for (void* p : {(void*)L.mask, L.S, L.lam,
                  L.W2})
    if (p) (void)hipFree(p);
  L = Level{};
}

// ---- coarse solve: PCG on the assembled vertex lattice ----------------------
//
// The coarse space (lx = 2) is continuous and its unknowns are the box's
// vertices, so the coarse PCG runs on the (nx+1)(ny+1)(nz+1) lattice with
// the assembled operator as 27 coefficients per vertex row: the operator
// application gs(ax_helm(p)) with the mask is one stencil pass, and the
// multiplicity-weighted inner products of the element-local form are plain
// sums over vertices.  The iteration is the oracle's (r = f, x = 0, Jacobi,
// p = z + beta p, crs_iter fixed iterations), in two kernels per iteration
// with no host synchronisation: the stencil pass (which forms p = z + beta p
// at the 27 neighbours itself, so p needs no pass of its own, and writes the
// partials of <w, p>) and the update (x, r, z and the partials of <r, z>).
// Every block sums the previous kernel's block partials itself, in one fixed
// order, so all blocks hold the same scalars and no finalise kernel runs:
// 1 + 2 crs_iter kernels where the element-local form took 57 (ax, gs,
// reductions, updates and masks).

// The sum of the n partials on every thread of the block (the same fixed
// order in every block).  sh: kHB / 64 + 1 doubles.
__device__ __forceinline__ double crs_all_sum(const double* __restrict__ part,
                                              int n, double* sh) {
  const double s = sum_partials<kHB>(part, n, sh);
  if (threadIdx.x == 0) sh[kHB / 64] = s;
  __syncthreads();
  const double t = sh[kHB / 64];
  __syncthreads();
  return t;
}

__device__ __forceinline__ bool lattice_masked(int faces, int nx, int ny, int nz,
                                               int a, int b, int c) {
  return ((faces & NEKOAMD_FACE_XMIN) && a == 0) ||
         ((faces & NEKOAMD_FACE_XMAX) && a == nx) ||
         ((faces & NEKOAMD_FACE_YMIN) && b == 0) ||
         ((faces & NEKOAMD_FACE_YMAX) && b == ny) ||
         ((faces & NEKOAMD_FACE_ZMIN) && c == 0) ||
         ((faces & NEKOAMD_FACE_ZMAX) && c == nz);
}

// r = mask(gs(r1)) on the lattice (the element-local copies of each vertex
// summed in element order), z = dinv r, partials of <r, z>.
__global__ void __launch_bounds__(kHB)
    crs_init_k(int nx, int ny, int nz, int faces, const double* __restrict__ r1,
               const double* __restrict__ dl, double* __restrict__ rl,
               double* __restrict__ zl, double* __restrict__ part) {
  __shared__ double sh[kHB / 64 + 1];
  const int nxl = nx + 1, nyl = ny + 1;
  const int64_t nl = (int64_t)nxl * nyl * (nz + 1);
  double s = 0.0;
  for (int64_t v = (int64_t)blockIdx.x * kHB + threadIdx.x; v < nl;
       v += (int64_t)gridDim.x * kHB) {
    const int a = (int)(v % nxl), b = (int)((v / nxl) % nyl),
              c = (int)(v / ((int64_t)nxl * nyl));
    double r = 0.0;
    if (!lattice_masked(faces, nx, ny, nz, a, b, c)) {
      for (int ez = c - 1; ez <= c; ++ez) {
        if (ez < 0 || ez >= nz) continue;
        for (int ey = b - 1; ey <= b; ++ey) {
          if (ey < 0 || ey >= ny) continue;
          for (int ex = a - 1; ex <= a; ++ex) {
            if (ex < 0 || ex >= nx) continue;
            const int64_t e = ex + (int64_t)nx * (ey + (int64_t)ny * ez);
            r += r1[e * 8 + (a - ex) + 2 * ((b - ey) + 2 * (c - ez))];
          }
        }
      }
    }
    const double z = r * dl[v];
    rl[v] = r;
    zl[v] = z;
    s += r * z;
  }
  s = block_sum<kHB>(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// p = z + beta p (beta = <r,z> / <r,z>_prev, 0 on the first iteration),
// w = A p by the 27-point rows, partials of <w, p>.  p is read from pold at
// the neighbours and written to pnew.  The 27 terms are unrolled with their
// loads issued together: a neighbour outside the lattice reads the vertex
// itself (always in bounds) and its term is dropped by a select, so the
// sum is the same as with the neighbours skipped.  Lattice indices are
// 32-bit (the set-up checks the vertex count).
template <bool FIRST>
__global__ void __launch_bounds__(kHB)
    crs_stencil_k(int nxl, int nyl, int nzl, const double* __restrict__ S,
                  const double* __restrict__ zl, const double* __restrict__ pold,
                  double* __restrict__ pnew, double* __restrict__ wl,
                  const double* __restrict__ rtz, const double* __restrict__ rtz_prev,
                  int G, double* __restrict__ part) {
  __shared__ double sh[kHB / 64 + 1];
  double beta = 0.0;
  if constexpr (!FIRST) {
    const double r1 = crs_all_sum(rtz, G, sh), r2 = crs_all_sum(rtz_prev, G, sh);
    beta = r2 != 0.0 ? r1 / r2 : 0.0;
  }
  const int nl = nxl * nyl * nzl, nxy = nxl * nyl;
  double s = 0.0;
  for (int v = blockIdx.x * kHB + threadIdx.x; v < nl; v += gridDim.x * kHB) {
    const int a = v % nxl, b = (v / nxl) % nyl, c = v / nxy;
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < 27; ++q) {
      const int di = q % 3 - 1, dj = (q / 3) % 3 - 1, dk = q / 9 - 1;
      const bool in = a + di >= 0 && a + di < nxl && b + dj >= 0 && b + dj < nyl &&
                      c + dk >= 0 && c + dk < nzl;
      const int m = in ? v + di + nxl * dj + nxy * dk : v;
      const double pm = FIRST ? zl[m] : zl[m] + beta * pold[m];
      const double t = S[q * nl + v] * pm;
      acc += in ? t : 0.0;
    }
    const double pv = FIRST ? zl[v] : zl[v] + beta * pold[v];
    pnew[v] = pv;
    wl[v] = acc;
    s += acc * pv;
  }
  s = block_sum<kHB>(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// alpha = <r,z> / <w,p>; x += alpha p (x = alpha p on the first iteration).
// Not last: r -= alpha w, z = dinv r, partials of <r, z>.  Last: x goes
// straight to the element-local e1 (every copy of each vertex), r is left.
template <bool FIRST, bool LAST>
__global__ void __launch_bounds__(kHB)
    crs_update_k(int nx, int ny, int nz, double* __restrict__ xl,
                 double* __restrict__ rl, double* __restrict__ zl,
                 const double* __restrict__ pl, const double* __restrict__ wl,
                 const double* __restrict__ dl, const double* __restrict__ pap,
                 const double* __restrict__ rtz, int G,
                 double* __restrict__ part, double* __restrict__ e1) {
  __shared__ double sh[kHB / 64 + 1];
  const double den = crs_all_sum(pap, G, sh), num = crs_all_sum(rtz, G, sh);
  const double alpha = den != 0.0 ? num / den : 0.0;
  const int nxl = nx + 1, nyl = ny + 1;
  const int64_t nl = (int64_t)nxl * nyl * (nz + 1);
  double s = 0.0;
  for (int64_t v = (int64_t)blockIdx.x * kHB + threadIdx.x; v < nl;
       v += (int64_t)gridDim.x * kHB) {
    const double x = (FIRST ? 0.0 : xl[v]) + alpha * pl[v];
    if constexpr (LAST) {
      const int a = (int)(v % nxl), b = (int)((v / nxl) % nyl),
                c = (int)(v / ((int64_t)nxl * nyl));
      for (int ez = c - 1; ez <= c; ++ez) {
        if (ez < 0 || ez >= nz) continue;
        for (int ey = b - 1; ey <= b; ++ey) {
          if (ey < 0 || ey >= ny) continue;
          for (int ex = a - 1; ex <= a; ++ex) {
            if (ex < 0 || ex >= nx) continue;
            const int64_t e = ex + (int64_t)nx * (ey + (int64_t)ny * ez);
            e1[e * 8 + (a - ex) + 2 * ((b - ey) + 2 * (c - ez))] = x;
          }
        }
      }
    } else {
      xl[v] = x;
      const double r = rl[v] - alpha * wl[v];
      const double z = r * dl[v];
      rl[v] = r;
      zl[v] = z;
      s += r * z;
    }
  }
  if constexpr (!LAST) {
    s = block_sum<kHB>(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
  }
}

// e1 = crs_iter PCG iterations from r1 (element-local, before its gs), the
// result masked and element-local: 1 + 2 crs_iter launches on the library
// stream (or recorded as kernel nodes).
int coarse_pcg(Hsmg* h) {
  const int G = h->crs_G, it = h->crs_iter;
  double* rtz = h->cpart;           // [it][G]: <r, z> before iteration q
  double* pap = h->cpart + (int64_t)it * G;  // [it][G]: <w, p> of iteration q
  nk_launch(crs_init_k, dim3(G), dim3(kHB), 0, h->nx, h->ny, h->nz, h->faces,
            (const double*)h->r1, (const double*)h->dl, h->rl, h->zl, rtz);
  NK_LAUNCHED();
  const int nxl = h->nx + 1, nyl = h->ny + 1, nzl = h->nz + 1;
  for (int q = 0; q < it; ++q) {
    double* pnew = h->pl[q & 1];
    const double* pold = h->pl[(q + 1) & 1];
    const double* rq = rtz + (int64_t)q * G;
    double* pq = pap + (int64_t)q * G;
    if (q == 0)
      nk_launch(crs_stencil_k<true>, dim3(G), dim3(kHB), 0, nxl, nyl, nzl,
                (const double*)h->S27, (const double*)h->zl, pold, pnew, h->wl,
                rq, rq, G, pq);
    else
      nk_launch(crs_stencil_k<false>, dim3(G), dim3(kHB), 0, nxl, nyl, nzl,
                (const double*)h->S27, (const double*)h->zl, pold, pnew, h->wl,
                rq, rq - G, G, pq);
    NK_LAUNCHED();
    double* rn = q + 1 < it ? rtz + (int64_t)(q + 1) * G : nullptr;
    auto* k = q + 1 == it ? (q == 0 ? crs_update_k<true, true> : crs_update_k<false, true>)
                          : (q == 0 ? crs_update_k<true, false> : crs_update_k<false, false>);
    nk_launch(k, dim3(G), dim3(kHB), 0, h->nx, h->ny, h->nz, h->xl, h->rl, h->zl,
              (const double*)pnew, (const double*)h->wl, (const double*)h->dl,
              (const double*)pq, rq, G, rn, h->e1);
    NK_LAUNCHED();
  }
  return kOk;
}

void hsmg_destroy(Hsmg* h) {
  for (Level& L : h->lv) level_free(L);
  for (void* p : {(void*)h->J_fm, (void*)h->J_mc, (void*)h->g[0], (void*)h->g[1],
                  (void*)h->g[2], (void*)h->g[3], (void*)h->g[4], (void*)h->g[5],
                  (void*)h->B, (void*)h->dx, (void*)h->h1, (void*)h->d, h->r2, h->e2,
                  (void*)h->r1, (void*)h->e1, h->rf, h->zf, (void*)h->S27,
                  (void*)h->dl, (void*)h->xl, (void*)h->rl, (void*)h->zl,
                  (void*)h->wl, (void*)h->pl[0], (void*)h->pl[1],
                  (void*)h->cpart})
    if (p) (void)hipFree(p);
  if (h->crs_graph) (void)hipGraphExecDestroy(h->crs_graph);
  if (h->side) {
    (void)hipStreamSynchronize(h->side);
    (void)hipStreamDestroy(h->side);
  }
  if (h->ev_r1) (void)hipEventDestroy(h->ev_r1);
  if (h->ev_crs) (void)hipEventDestroy(h->ev_crs);
  delete h;
}

// The coarse solve as a HIP graph only with NEKOAMD_HIP_GRAPHS=1 (read when a
// multigrid object is created); by default its kernels are queued one by one
// on the side stream.  The graph form is also the one the kernel tracer
// (rocprofv3 --kernel-trace) faults under, in its graph launch path.
static bool graphs_requested() {
  const char* v = std::getenv("NEKOAMD_HIP_GRAPHS");
  return v && v[0] == '1';
}

// e1 = crs_iter PCG iterations on the coarse level from r1, on the side
// stream (default priority) once r1 is ready, so that they run under the
// fine and mid Schwarz steps; ev_crs is recorded when they are done and the
// library stream waits on it before reading e1 (coarse_join).  Without
// graphs the 1 + 2 crs_iter kernels are queued on the side stream; with
// graphs they are one graph launch there, the graph built on the first call
// by recording coarse_pcg's launches as kernel nodes (LaunchRecorder), not
// by stream capture: nothing is queued on any stream while it is built.
int coarse_solve(Hsmg* h) {
  if (!h->side) {
    NK_HIP(hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking));
    NK_HIP(hipEventCreateWithFlags(&h->ev_r1, hipEventDisableTiming));
    NK_HIP(hipEventCreateWithFlags(&h->ev_crs, hipEventDisableTiming));
  }
  if (h->graphs && !h->crs_graph) {
    LaunchRecorder rec;
    NK_HIP(hipGraphCreate(&rec.graph, 0));
    launch_recorder() = &rec;
    const int rc = coarse_pcg(h);
    launch_recorder() = nullptr;
    hipError_t e = rec.err;
    if (rc == kOk && e == hipSuccess)
      e = hipGraphInstantiate(&h->crs_graph, rec.graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(rec.graph);
    if (rc != kOk) return rc;
    if (e != hipSuccess) {
      h->crs_graph = nullptr;
      return fail_hip("hsmg_apply", e);
    }
  }
  NK_HIP(hipEventRecord(h->ev_r1, stream()));
  NK_HIP(hipStreamWaitEvent(h->side, h->ev_r1, 0));
  if (h->graphs) {
    NK_HIP(hipGraphLaunch(h->crs_graph, h->side));
  } else {
    launch_stream() = h->side;
    const int rc = coarse_pcg(h);
    launch_stream() = nullptr;
    if (rc != kOk) return rc;
  }
  NK_HIP(hipEventRecord(h->ev_crs, h->side));
  return kOk;
}

// The library stream waits for the coarse solve.
int coarse_join(Hsmg* h) {
  if (h->ev_crs) NK_HIP(hipStreamWaitEvent(stream(), h->ev_crs, 0));
  return kOk;
}

// The lattice form of the coarse operator: each element's 8 x 8 matrix from
// ax_helm (lx = 2, h1 = 1) applied to the 8 element-local unit fields,
// summed into the 27 coefficients of each vertex row in element order (rows
// of masked vertices left zero); the inverse diagonal from the element-local
// Jacobi d (zero on masked vertices); the solve's buffers.
int coarse_lattice_setup(Hsmg* h) {
  Level& L1 = h->lv[0];
  const int ne = h->nelv, nx = h->nx, ny = h->ny, nz = h->nz;
  const int nxl = nx + 1, nyl = ny + 1;
  const int64_t nl = (int64_t)nxl * nyl * (nz + 1), n1 = L1.n;
  if (27 * nl >= ((int64_t)1 << 31))
    return fail_arg("nekoamd_hsmg_create", "coarse lattice too large");
  h->nl = nl;
  h->crs_G = nk_grid(nl, kHB, 512);
  std::vector<double> A((size_t)ne * 64), col(n1), d(n1);
  double *u = nullptr, *w = nullptr;
  hipError_t e = hipMalloc(&u, sizeof(double) * n1);
  if (e == hipSuccess) e = hipMalloc(&w, sizeof(double) * n1);
  int rc = kOk;
  for (int c = 0; c < 8 && e == hipSuccess && rc == kOk; ++c) {
    for (int64_t l = 0; l < n1; ++l) col[l] = (l % 8) == c ? 1.0 : 0.0;
    e = hipMemcpy(u, col.data(), sizeof(double) * n1, hipMemcpyHostToDevice);
    if (e == hipSuccess)
      rc = ax_helm_launch<double, double>(w, u, h->dx, h->h1, nullptr, h->g[0],
                                          h->g[1], h->g[2], h->g[3], h->g[4],
                                          h->g[5], h->B, 2, ne, false, nullptr);
    if (e == hipSuccess && rc == kOk) e = hipStreamSynchronize(stream());
    if (e == hipSuccess && rc == kOk)
      e = hipMemcpy(col.data(), w, sizeof(double) * n1, hipMemcpyDeviceToHost);
    if (e == hipSuccess && rc == kOk)
      for (int el = 0; el < ne; ++el)
        for (int l = 0; l < 8; ++l) A[(size_t)el * 64 + l * 8 + c] = col[(size_t)el * 8 + l];
  }
  if (e == hipSuccess && rc == kOk) e = hipStreamSynchronize(stream());
  if (e == hipSuccess && rc == kOk)
    e = hipMemcpy(d.data(), h->d, sizeof(double) * n1, hipMemcpyDeviceToHost);
  (void)hipFree(u);
  (void)hipFree(w);
  if (rc != kOk) return rc;
  if (e != hipSuccess) return fail_hip("nekoamd_hsmg_create", e);
  auto masked = [&](int a, int b, int c) {
    const int f = h->faces;
    return ((f & NEKOAMD_FACE_XMIN) && a == 0) || ((f & NEKOAMD_FACE_XMAX) && a == nx) ||
           ((f & NEKOAMD_FACE_YMIN) && b == 0) || ((f & NEKOAMD_FACE_YMAX) && b == ny) ||
           ((f & NEKOAMD_FACE_ZMIN) && c == 0) || ((f & NEKOAMD_FACE_ZMAX) && c == nz);
  };
  std::vector<double> S((size_t)27 * nl, 0.0), dl(nl, 0.0);
  for (int ez = 0; ez < nz; ++ez)
    for (int ey = 0; ey < ny; ++ey)
      for (int ex = 0; ex < nx; ++ex) {
        const int64_t el = ex + (int64_t)nx * (ey + (int64_t)ny * ez);
        for (int l = 0; l < 8; ++l) {
          const int i = l & 1, j = (l >> 1) & 1, k = l >> 2;
          const int a = ex + i, b = ey + j, c = ez + k;
          if (masked(a, b, c)) continue;
          const int64_t v = a + (int64_t)nxl * (b + (int64_t)nyl * c);
          dl[v] = d[(size_t)el * 8 + l];
          for (int m = 0; m < 8; ++m) {
            const int di = (m & 1) - i, dj = ((m >> 1) & 1) - j, dk = (m >> 2) - k;
            const int q = (di + 1) + 3 * ((dj + 1) + 3 * (dk + 1));
            S[(size_t)q * nl + v] += A[(size_t)el * 64 + l * 8 + m];
          }
        }
      }
  const size_t vb = sizeof(double) * (size_t)nl;
  e = hipMalloc(&h->S27, 27 * vb);
  for (double** p : {&h->dl, &h->xl, &h->rl, &h->zl, &h->wl, &h->pl[0], &h->pl[1]})
    if (e == hipSuccess) e = hipMalloc(p, vb);
  if (e == hipSuccess)
    e = hipMalloc(&h->cpart, sizeof(double) * 2 * (size_t)h->crs_iter * h->crs_G);
  if (e == hipSuccess) e = hipMemcpy(h->S27, S.data(), 27 * vb, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(h->dl, dl.data(), vb, hipMemcpyHostToDevice);
  if (e != hipSuccess) return fail_hip("nekoamd_hsmg_create", e);
  return kOk;
}

}  // namespace

namespace {

template <typename T>
int apply_t(Hsmg* h, double* z, const double* r) {
  Level &L3 = h->lv[2], &L2 = h->lv[1];
  const int ne = h->nelv;
  T* zf = (T*)h->zf;
  T* r2 = (T*)h->r2;
  T* e2 = (T*)h->e2;
  int rc;
  // the fine residual in the data-path type
  const T* rf = nullptr;
  if constexpr (sizeof(T) == sizeof(double)) {
    rf = (const T*)r;
  } else {
    hipLaunchKernelGGL((hs_cast_k<T, double>), dim3(nk_grid(L3.n, kHB)), dim3(kHB), 0,
                       stream(), L3.n, (T*)h->rf, r);
    NK_LAUNCHED();
    rf = (const T*)h->rf;
  }
  // restrict r to mid and on to coarse (FP64) first: the coarse solve needs
  // only r1, so it runs (on the side stream) under the two Schwarz steps
  if ((rc = restrict_launch<T, T>(h, L3.lx, L2.lx, r2, rf, h->J_fm)) != kOk)
    return rc;
  if ((rc = gs_launch<T>(L2.gs, r2, 1, nullptr)) != kOk) return rc;
  // (the coarse gs and mask are taken by the lattice solve's first kernel)
  if ((rc = restrict_launch<T, double>(h, L2.lx, 2, h->r1, r2, h->J_mc)) != kOk)
    return rc;
  if ((rc = coarse_solve(h)) != kOk) return rc;
  // fine Schwarz into a private buffer (masked and weighted), then the mid
  // Schwarz (r2 is only read)
  if ((rc = schwarz<T>(h, L3, zf, rf, true)) != kOk) return rc;
  if ((rc = schwarz<T>(h, L2, e2, r2, true)) != kOk) return rc;
  if ((rc = coarse_join(h)) != kOk) return rc;
  // prolongate and combine: e2 += J_mc e1, then z = zf + J_fm e2 in one
  // pass.  Both terms are continuous (zf is gs-assembled, the weights equal
  // on every copy of a point, and an element-local interpolation of a
  // continuous field takes its face values from the shared face alone), so
  // the closing mult gs of the method, z = mult gs(zf + J_fm e2), is the
  // identity up to rounding and is left out.
  if ((rc = prolong_launch<double, T, T>(2, L2.lx, ne, e2, e2, (const T*)nullptr,
                                         h->e1, h->J_mc)) != kOk)
    return rc;
  if ((rc = prolong_launch<T, double, T>(L2.lx, L3.lx, ne, z, zf, (const T*)nullptr, e2,
                                         h->J_fm)) != kOk)
    return rc;
  return kOk;
}

}  // namespace

int hsmg_apply(void* hp, double* z, const double* r) {
  auto* h = static_cast<Hsmg*>(hp);
  NK_REQUIRE(h != nullptr, "null hsmg");
  return h->f32 ? apply_t<float>(h, z, r) : apply_t<double>(h, z, r);
}

}  // namespace nekoamd

using namespace nekoamd;

extern "C" {

int nekoamd_hsmg_create(void** st, const int* nx, const int* ny, const int* nz,
                        const int* lx, const double* corners,
                        const int* dirichlet_faces, const int* crs_iter,
                        const int* precision) {
  NK_INIT();
  NK_REQUIRE(st && nx && ny && nz && lx && corners && dirichlet_faces && crs_iter &&
                 precision,
             "null argument");
  NK_REQUIRE(*precision == NEKOAMD_CG_FP64 || *precision == NEKOAMD_CG_MIXED,
             "precision must be NEKOAMD_CG_FP64 or NEKOAMD_CG_MIXED");
  NK_REQUIRE(*nx >= 1 && *ny >= 1 && *nz >= 1, "empty mesh");
  if (!(*lx == 4 || *lx == 6 || *lx == 8)) {
    set_error(__func__, "hsmg supports lx in {4, 6, 8}");
    return kErrUnsupported;
  }
  NK_REQUIRE(*crs_iter >= 1, "crs_iter < 1");
  auto* h = new Hsmg;
  h->graphs = graphs_requested();
  h->nx = *nx;
  h->ny = *ny;
  h->nz = *nz;
  h->nelv = *nx * *ny * *nz;
  h->lx = *lx;
  h->lxm = *lx >= 5 ? 4 : std::max(*lx - 1, 3);
  h->faces = *dirichlet_faces;
  h->crs_iter = *crs_iter;
  h->f32 = *precision == NEKOAMD_CG_MIXED;
  const int ne = h->nelv;
  // element lengths: mean of the four edges per direction
  std::vector<double> len((size_t)ne * 3, 0.0);
  for (int e = 0; e < ne; ++e)
    for (int d = 0; d < 3; ++d) {
      double acc = 0.0;
      for (int k = 0; k < 8; ++k) {
        if (k & (1 << d)) continue;
        const double* a = corners + ((size_t)e * 8 + k) * 3;
        const double* b = corners + ((size_t)e * 8 + (k | (1 << d))) * 3;
        acc += std::sqrt((b[0] - a[0]) * (b[0] - a[0]) + (b[1] - a[1]) * (b[1] - a[1]) +
                         (b[2] - a[2]) * (b[2] - a[2]));
      }
      len[(size_t)e * 3 + d] = acc / 4.0;
    }
  int rc = kOk;
  auto fail = [&](int code) {
    hsmg_destroy(h);
    return code;
  };
  if ((rc = level_setup(h, h->lv[2], h->lx, true, len)) != kOk) return fail(rc);
  if ((rc = level_setup(h, h->lv[1], h->lxm, true, len)) != kOk) return fail(rc);
  if ((rc = level_setup(h, h->lv[0], 2, false, len)) != kOk) return fail(rc);
  // interpolation matrices
  std::vector<double> w3, A3, z3, w2, A2, z2, w1, A1, z1;
  one_d(h->lx, w3, A3, z3);
  one_d(h->lxm, w2, A2, z2);
  one_d(2, w1, A1, z1);
  auto lag = [](const std::vector<double>& zf, const std::vector<double>& zc) {
    std::vector<double> J(zf.size() * zc.size());
    for (size_t q = 0; q < zf.size(); ++q)
      for (size_t i = 0; i < zc.size(); ++i) {
        double v = 1.0;
        for (size_t k = 0; k < zc.size(); ++k)
          if (k != i) v *= (zf[q] - zc[k]) / (zc[i] - zc[k]);
        J[q * zc.size() + i] = v;
      }
    return J;
  };
  const std::vector<double> Jfm = lag(z3, z2), Jmc = lag(z2, z1);
  hipError_t e = hipMalloc(&h->J_fm, sizeof(double) * Jfm.size());
  if (e == hipSuccess) e = hipMalloc(&h->J_mc, sizeof(double) * Jmc.size());
  if (e == hipSuccess)
    e = hipMemcpy(h->J_fm, Jfm.data(), sizeof(double) * Jfm.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(h->J_mc, Jmc.data(), sizeof(double) * Jmc.size(), hipMemcpyHostToDevice);
  // coarse operator: geometric factors of the lx = 2 space, Jacobi, PCG
  Level& L1 = h->lv[0];
  const int two = 2;
  double* cd = nullptr;
  double *zd = nullptr, *wd = nullptr;
  std::vector<double> D1(4);
  {
    std::vector<double> zz(2), ww(2);
    nekoamd_gll(&two, zz.data(), ww.data(), D1.data());
  }
  for (int k = 0; k < 6 && e == hipSuccess; ++k) e = hipMalloc(&h->g[k], sizeof(double) * L1.n);
  for (double** p : {&h->B, &h->h1, &h->d, &h->r1, &h->e1})
    if (e == hipSuccess) e = hipMalloc(p, sizeof(double) * L1.n);
  if (e == hipSuccess) e = hipMalloc(&h->dx, sizeof(double) * 4);
  if (e == hipSuccess) e = hipMalloc(&cd, sizeof(double) * ne * 24);
  if (e == hipSuccess) e = hipMalloc(&zd, sizeof(double) * 2);
  if (e == hipSuccess) e = hipMalloc(&wd, sizeof(double) * 2);
  if (e == hipSuccess) e = hipMemcpy(h->dx, D1.data(), sizeof(double) * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(cd, corners, sizeof(double) * ne * 24, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(zd, z1.data(), sizeof(double) * 2, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(wd, w1.data(), sizeof(double) * 2, hipMemcpyHostToDevice);
  const size_t es = h->f32 ? sizeof(float) : sizeof(double);
  for (void** p : {&h->r2, &h->e2})
    if (e == hipSuccess) e = hipMalloc(p, es * h->lv[1].n);
  if (e == hipSuccess) e = hipMalloc(&h->zf, es * h->lv[2].n);
  if (e == hipSuccess && h->f32) e = hipMalloc(&h->rf, es * h->lv[2].n);
  if (e != hipSuccess) {
    (void)hipFree(cd);
    (void)hipFree(zd);
    (void)hipFree(wd);
    return fail(fail_hip(__func__, e));
  }
  rc = nekoamd_coef_geo(nullptr, nullptr, nullptr, h->g[0], h->g[1], h->g[2], h->g[3],
                        h->g[4], h->g[5], h->B, nullptr, cd, zd, wd, h->dx, &two, &h->nelv);
  (void)hipStreamSynchronize(stream());
  (void)hipFree(cd);
  (void)hipFree(zd);
  (void)hipFree(wd);
  if (rc != kOk) return fail(rc);
  const double one = 1.0;
  const int ifh2 = 0;
  if ((rc = nekoamd_cfill(h->h1, &one, &L1.n)) != kOk) return fail(rc);
  if ((rc = nekoamd_jacobi_update(h->d, L1.gs, h->dx, h->g[0], h->g[3], h->g[5], h->h1,
                                  nullptr, h->B, &two, &h->nelv, &ifh2)) != kOk)
    return fail(rc);
  if ((rc = coarse_lattice_setup(h)) != kOk) return fail(rc);
  NK_HIP(hipStreamSynchronize(stream()));
  *st = h;
  return kOk;
}

int nekoamd_hsmg_free(void* st) {
  NK_INIT();
  if (!st) return kOk;
  NK_HIP(hipStreamSynchronize(stream()));
  hsmg_destroy(static_cast<Hsmg*>(st));
  return kOk;
}

int nekoamd_hsmg_apply(void* st, void* z, const void* r) {
  NK_INIT();
  NK_REQUIRE(st && z && r, "null argument");
  return hsmg_apply(st, (double*)z, (const double*)r);
}

}  // extern "C"
