// gs.hip -- gather-scatter (direct stiffness summation, Q Q^T) on gfx950.
//
// Replaces Neko's gs_t as the reference uses it: init(dofmap) at
// sources/adjoint/adjoint_scheme.f90:341, op(field, GS_OP_ADD) at
// sources/adjoint/adjoint_pnpn.f90:725, :755-757 and inside every Krylov
// iteration, free at adjoint_scheme.f90:913.
//
// Map layout.  Only points shared by m >= 2 elements are touched.  They are
// stored as tuples of m 0-based int32 local indices, bucketed by m, so a
// thread owns one global point: it loads its m indices with one or two
// 16-byte loads, gathers, combines the m values in member order and writes
// the result back to all m copies.  No atomics, no segment offsets, and the
// combination order is fixed, so the result is bit-identical to a
// sequential CPU pass over the same map.  Each bucket starts on a 16-byte
// boundary in device memory.
//
// Launch layout of the 2-, 4- and 8-tuples (faces, edges and vertices of a
// hexahedral mesh; one launch for all three, gs_blk_k).  In the element
// layout every 128-byte line of a field holds points of some tuple (the
// x-face members sit at i = 0 and i = lx-1 of every row), so the least a
// gather-scatter can move is each line read once and written once.  It gets
// there when all tuples that touch a line run close in time on the XCD whose
// L2 holds it: tuples are grouped by their smallest member's index (runs of
// kBlkGroup local points, i.e. 16 elements at lx = 8), each group's 2-, 4-
// and 8-tuples in blocks of up to 256, and the block table is interleaved so
// that workgroup b, which the dispatcher sends to XCD b mod nxcd, takes the
// logical blocks of one contiguous eighth of the points.  At 32^3 elements,
// p = 7, a 4 MB 16-way LRU model per XCD gives 2.69x the algorithmic bytes
// for the tuples in id order (the measured figure) and 1.61x for this layout.
#include <algorithm>
#include <cstring>
#include <vector>

#include "internal.h"
#include "nekoamd.h"

namespace nekoamd {

int build_gs_map(int64_t n, const int64_t* gid, std::vector<int>& bm,
                 std::vector<int64_t>& bcnt, std::vector<int32_t>& idx);

struct GsMap {
  int64_t n = 0;
  int nb = 0;
  int m[NEKOAMD_GS_MAX_MULT];
  int64_t cnt[NEKOAMD_GS_MAX_MULT];
  int64_t host_off[NEKOAMD_GS_MAX_MULT];  // offsets in the canonical map
  int64_t dev_off[NEKOAMD_GS_MAX_MULT];   // padded offsets on the device
  int64_t nocc = 0;
  int64_t dev_len = 0;
  int32_t* idx_d = nullptr;
  // launch layout of the 2-, 4- and 8-tuples: per workgroup {m, offset / 4
  // into blk_idx, count, index of its first tuple}; empty entries pad the
  // table to a multiple of the XCD count
  int32_t* blk_idx = nullptr;
  int4* blk_desc = nullptr;
  int nblk = 0;
  int64_t blk_len = 0, blk_tuples = 0;
  // mass average in the launch layout: B per occurrence, 1/gs(B) per tuple
  double* avg_bb = nullptr;
  double* avg_binvb = nullptr;
  // mass average: B per occurrence in the layout of idx_d and 1/gs(B) per
  // group, gathered on the first call with a given (B, Binv) (keyed by their
  // addresses: the mass of a mesh is fixed once built)
  double* avg_b = nullptr;     // dev_len
  double* avg_binv = nullptr;  // one per group, buckets in order
  int64_t goff[NEKOAMD_GS_MAX_MULT] = {};
  const void* avg_key[2] = {nullptr, nullptr};
};

template <int OP, typename T>
__device__ __forceinline__ T gs_combine(T s, T v) {
  if constexpr (OP == NEKOAMD_GS_ADD) return s + v;
  if constexpr (OP == NEKOAMD_GS_MUL) return s * v;
  if constexpr (OP == NEKOAMD_GS_MIN) return v < s ? v : s;
  return v > s ? v : s;
}

// One group of fixed multiplicity M (2, 4, 8): vector index loads,
// unrolled.
template <typename T, int M, int OP>
__device__ __forceinline__ void gs_group(T* __restrict__ u,
                                         const int32_t* __restrict__ idx,
                                         int64_t t) {
  int32_t q[M];
  if constexpr (M == 2) {
    const int2 a = reinterpret_cast<const int2*>(idx)[t];
    q[0] = a.x;
    q[1] = a.y;
  } else {
#pragma unroll
    for (int c = 0; c < M / 4; ++c) {
      const int4 a = reinterpret_cast<const int4*>(idx)[t * (M / 4) + c];
      q[4 * c + 0] = a.x;
      q[4 * c + 1] = a.y;
      q[4 * c + 2] = a.z;
      q[4 * c + 3] = a.w;
    }
  }
  T v[M];
#pragma unroll
  for (int r = 0; r < M; ++r) v[r] = u[q[r]];
  T s = v[0];
#pragma unroll
  for (int r = 1; r < M; ++r) s = gs_combine<OP>(s, v[r]);
#pragma unroll
  for (int r = 0; r < M; ++r) u[q[r]] = s;
}

template <typename T, int M, int OP>
__global__ void __launch_bounds__(256)
    gs_fixed_k(T* __restrict__ u, const int32_t* __restrict__ idx,
               int64_t cnt, const int* __restrict__ skip) {
  if (skip != nullptr && *skip) return;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt) return;
  gs_group<T, M, OP>(u, idx, t);
}

// Workgroup b takes entry b of the block table: up to 256 tuples of one
// multiplicity.
template <typename T, int OP>
__global__ void __launch_bounds__(256)
    gs_blk_k(T* __restrict__ u, const int32_t* __restrict__ bidx,
             const int4* __restrict__ desc, const int* __restrict__ skip) {
  if (skip != nullptr && *skip) return;
  const int4 d = desc[blockIdx.x];
  const int t = threadIdx.x;
  if (t >= d.z) return;
  const int32_t* idx = bidx + (int64_t)d.y * 4;
  if (d.x == 2)
    gs_group<T, 2, OP>(u, idx, t);
  else if (d.x == 4)
    gs_group<T, 4, OP>(u, idx, t);
  else
    gs_group<T, 8, OP>(u, idx, t);
}

// Three fields on the launch layout at once: each tuple's indices are
// loaded once and the tuple is combined in every field in turn (field c left
// alone if !active[c] or while *skip[c] is set), the same arithmetic per
// field as gs_blk_k.
template <typename T>
struct Gs3Vec {
  T* u[3];
  const int* skip[3];
  int active[3];
};

template <typename T, int M, int OP>
__device__ __forceinline__ void gs_group3(const Gs3Vec<T>& v, const bool* act,
                                          const int32_t* __restrict__ idx,
                                          int64_t t) {
  int32_t q[M];
  if constexpr (M == 2) {
    const int2 a = reinterpret_cast<const int2*>(idx)[t];
    q[0] = a.x;
    q[1] = a.y;
  } else {
#pragma unroll
    for (int c = 0; c < M / 4; ++c) {
      const int4 a = reinterpret_cast<const int4*>(idx)[t * (M / 4) + c];
      q[4 * c + 0] = a.x;
      q[4 * c + 1] = a.y;
      q[4 * c + 2] = a.z;
      q[4 * c + 3] = a.w;
    }
  }
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    if (!act[f]) continue;
    T* __restrict__ u = v.u[f];
    T x[M];
#pragma unroll
    for (int r = 0; r < M; ++r) x[r] = u[q[r]];
    T s = x[0];
#pragma unroll
    for (int r = 1; r < M; ++r) s = gs_combine<OP>(s, x[r]);
#pragma unroll
    for (int r = 0; r < M; ++r) u[q[r]] = s;
  }
}

template <typename T, int OP>
__global__ void __launch_bounds__(256)
    gs_blk3_k(Gs3Vec<T> v, const int32_t* __restrict__ bidx,
              const int4* __restrict__ desc) {
  bool act[3];
#pragma unroll
  for (int f = 0; f < 3; ++f)
    act[f] = v.active[f] && !(v.skip[f] != nullptr && *v.skip[f]);
  if (!act[0] && !act[1] && !act[2]) return;
  const int4 d = desc[blockIdx.x];
  const int t = threadIdx.x;
  if (t >= d.z) return;
  const int32_t* idx = bidx + (int64_t)d.y * 4;
  if (d.x == 2)
    gs_group3<T, 2, OP>(v, act, idx, t);
  else if (d.x == 4)
    gs_group3<T, 4, OP>(v, act, idx, t);
  else
    gs_group3<T, 8, OP>(v, act, idx, t);
}

// The buckets of multiplicity 2, 4 and 8 (faces, edges and vertices of a
// hexahedral mesh) in one launch: the first c2 threads take the 2-groups,
// the next c4 the 4-groups, the rest the 8-groups.  Every local point is in
// one group, so the buckets do not interact.
template <typename T, int OP>
__global__ void __launch_bounds__(256)
    gs_248_k(T* __restrict__ u, const int32_t* __restrict__ idx2, int64_t c2,
             const int32_t* __restrict__ idx4, int64_t c4,
             const int32_t* __restrict__ idx8, int64_t c8,
             const int* __restrict__ skip) {
  if (skip != nullptr && *skip) return;
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < c2) {
    gs_group<T, 2, OP>(u, idx2, t);
    return;
  }
  t -= c2;
  if (t < c4) {
    gs_group<T, 4, OP>(u, idx4, t);
    return;
  }
  t -= c4;
  if (t < c8) gs_group<T, 8, OP>(u, idx8, t);
}

// Mass average a = gs(B a) Binv (Binv = 1 / gs(B)) over the shared points:
// s = sum of B u over a group's members in member order, then u = s Binv at
// every member -- the products, sum and order of col2(B), gs ADD,
// col2(Binv).  B comes per occurrence in the map's order (bq, next to the
// indices) and Binv once per group (its value at every member), so only u
// is gathered.  A point in no group keeps u, the exact value there (the
// three full passes would round (u B) Binv).
template <int M>
__device__ __forceinline__ void gs_avg_group(double* __restrict__ u,
                                             const double* __restrict__ bq,
                                             const double* __restrict__ binv,
                                             const int32_t* __restrict__ idx,
                                             int64_t t) {
  int32_t q[M];
  double b[M];
  if constexpr (M == 2) {
    const int2 a = reinterpret_cast<const int2*>(idx)[t];
    q[0] = a.x;
    q[1] = a.y;
    const double2 w = reinterpret_cast<const double2*>(bq)[t];
    b[0] = w.x;
    b[1] = w.y;
  } else {
#pragma unroll
    for (int c = 0; c < M / 4; ++c) {
      const int4 a = reinterpret_cast<const int4*>(idx)[t * (M / 4) + c];
      q[4 * c + 0] = a.x;
      q[4 * c + 1] = a.y;
      q[4 * c + 2] = a.z;
      q[4 * c + 3] = a.w;
    }
#pragma unroll
    for (int c = 0; c < M / 2; ++c) {
      const double2 w = reinterpret_cast<const double2*>(bq)[t * (M / 2) + c];
      b[2 * c + 0] = w.x;
      b[2 * c + 1] = w.y;
    }
  }
  double s = u[q[0]] * b[0];
#pragma unroll
  for (int r = 1; r < M; ++r) s += u[q[r]] * b[r];
  const double v = s * binv[t];
#pragma unroll
  for (int r = 0; r < M; ++r) u[q[r]] = v;
}

struct AvgBucket {
  const int32_t* idx;
  const double* bq;
  const double* binv;
  int64_t cnt;
};

__global__ void __launch_bounds__(256)
    gs_avg_248_k(double* __restrict__ u, AvgBucket b2, AvgBucket b4,
                 AvgBucket b8) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < b2.cnt) {
    gs_avg_group<2>(u, b2.bq, b2.binv, b2.idx, t);
    return;
  }
  t -= b2.cnt;
  if (t < b4.cnt) {
    gs_avg_group<4>(u, b4.bq, b4.binv, b4.idx, t);
    return;
  }
  t -= b4.cnt;
  if (t < b8.cnt) gs_avg_group<8>(u, b8.bq, b8.binv, b8.idx, t);
}

__global__ void __launch_bounds__(256)
    gs_avg_any_k(double* __restrict__ u, const double* __restrict__ bq,
                 const double* __restrict__ binv,
                 const int32_t* __restrict__ idx, int64_t cnt, int m) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt) return;
  const int32_t* q = idx + t * m;
  const double* b = bq + t * m;
  double s = u[q[0]] * b[0];
  for (int r = 1; r < m; ++r) s += u[q[r]] * b[r];
  const double v = s * binv[t];
  for (int r = 0; r < m; ++r) u[q[r]] = v;
}

// bq[k] = B[idx[k]] over the device index array (padding reads B[0]);
// binv[g] = Binv at the first member of group g of a bucket.
// The mass average on the launch layout (see gs_blk_k).
__global__ void __launch_bounds__(256)
    gs_avg_blk_k(double* __restrict__ u, const int32_t* __restrict__ bidx,
                 const double* __restrict__ bq,
                 const double* __restrict__ binv,
                 const int4* __restrict__ desc) {
  const int4 d = desc[blockIdx.x];
  const int t = threadIdx.x;
  if (t >= d.z) return;
  const int64_t o = (int64_t)d.y * 4;
  if (d.x == 2)
    gs_avg_group<2>(u, bq + o, binv + d.w, bidx + o, t);
  else if (d.x == 4)
    gs_avg_group<4>(u, bq + o, binv + d.w, bidx + o, t);
  else
    gs_avg_group<8>(u, bq + o, binv + d.w, bidx + o, t);
}

// 1/gs(B) of each tuple of the launch layout, from its first member.
__global__ void __launch_bounds__(256)
    gs_avg_gather_binv_blk_k(const int32_t* __restrict__ bidx,
                             const int4* __restrict__ desc,
                             const double* __restrict__ Binv,
                             double* __restrict__ binv) {
  const int4 d = desc[blockIdx.x];
  const int t = threadIdx.x;
  if (t >= d.z) return;
  binv[d.w + t] = Binv[bidx[(int64_t)d.y * 4 + (int64_t)t * d.x]];
}

__global__ void __launch_bounds__(256)
    gs_avg_gather_b_k(int64_t len, const int32_t* __restrict__ idx,
                      const double* __restrict__ B, double* __restrict__ bq) {
  for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < len;
       k += (int64_t)gridDim.x * 256)
    bq[k] = B[idx[k]];
}

__global__ void __launch_bounds__(256)
    gs_avg_gather_binv_k(int64_t cnt, int m, const int32_t* __restrict__ idx,
                         const double* __restrict__ Binv,
                         double* __restrict__ binv) {
  for (int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x; g < cnt;
       g += (int64_t)gridDim.x * 256)
    binv[g] = Binv[idx[g * m]];
}

// Any multiplicity.
template <typename T, int OP>
__global__ void __launch_bounds__(256)
    gs_any_k(T* __restrict__ u, const int32_t* __restrict__ idx, int64_t cnt,
             int m, const int* __restrict__ skip) {
  if (skip != nullptr && *skip) return;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt) return;
  const int32_t* q = idx + t * m;
  T s = u[q[0]];
  for (int r = 1; r < m; ++r) s = gs_combine<OP>(s, u[q[r]]);
  for (int r = 0; r < m; ++r) u[q[r]] = s;
}

template <typename T, int OP>
static void gs_bucket(T* u, const GsMap* g, int b, const int* skip) {
  const int64_t cnt = g->cnt[b];
  const int32_t* idx = g->idx_d + g->dev_off[b];
  const int blocks = (int)((cnt + 255) / 256);
  switch (g->m[b]) {
    case 2:
      nk_launch(gs_fixed_k<T, 2, OP>, dim3(blocks), dim3(256), 0, u, idx, cnt,
                skip);
      break;
    case 4:
      nk_launch(gs_fixed_k<T, 4, OP>, dim3(blocks), dim3(256), 0, u, idx, cnt,
                skip);
      break;
    case 8:
      nk_launch(gs_fixed_k<T, 8, OP>, dim3(blocks), dim3(256), 0, u, idx, cnt,
                skip);
      break;
    default:
      nk_launch(gs_any_k<T, OP>, dim3(blocks), dim3(256), 0, u, idx, cnt,
                g->m[b], skip);
  }
}

template <typename T, int OP>
static void gs_248(T* u, const GsMap* g, const int* skip) {
  if (g->nblk > 0) {
    nk_launch(gs_blk_k<T, OP>, dim3(g->nblk), dim3(256), 0, u, g->blk_idx,
              g->blk_desc, skip);
    return;
  }
  const int32_t* ix[3] = {nullptr, nullptr, nullptr};
  int64_t c[3] = {0, 0, 0};
  for (int b = 0; b < g->nb; ++b) {
    const int slot = g->m[b] == 2 ? 0 : g->m[b] == 4 ? 1 : g->m[b] == 8 ? 2 : -1;
    if (slot < 0) continue;
    ix[slot] = g->idx_d + g->dev_off[b];
    c[slot] = g->cnt[b];
  }
  const int64_t tot = c[0] + c[1] + c[2];
  if (tot == 0) return;
  nk_launch(gs_248_k<T, OP>, dim3((unsigned)((tot + 255) / 256)), dim3(256),
            0, u, ix[0], c[0], ix[1], c[1], ix[2], c[2], skip);
}

template <typename T>
int gs_launch(void* gs, T* u, int op, const int* skip) {
  const GsMap* g = (const GsMap*)gs;
  if (!g) return fail_arg("gs_op", "null gs handle");
  if (!u && g->n > 0) return fail_arg("gs_op", "null field");
  // multiplicities 2, 4, 8 in one launch, any other bucket on its own
  switch (op) {
    case NEKOAMD_GS_ADD: gs_248<T, NEKOAMD_GS_ADD>(u, g, skip); break;
    case NEKOAMD_GS_MUL: gs_248<T, NEKOAMD_GS_MUL>(u, g, skip); break;
    case NEKOAMD_GS_MIN: gs_248<T, NEKOAMD_GS_MIN>(u, g, skip); break;
    case NEKOAMD_GS_MAX: gs_248<T, NEKOAMD_GS_MAX>(u, g, skip); break;
    default: return fail_arg("gs_op", "unknown gs operation");
  }
  NK_LAUNCHED();
  for (int b = 0; b < g->nb; ++b) {
    if (g->cnt[b] == 0 || g->m[b] == 2 || g->m[b] == 4 || g->m[b] == 8)
      continue;
    switch (op) {
      case NEKOAMD_GS_ADD: gs_bucket<T, NEKOAMD_GS_ADD>(u, g, b, skip); break;
      case NEKOAMD_GS_MUL: gs_bucket<T, NEKOAMD_GS_MUL>(u, g, b, skip); break;
      case NEKOAMD_GS_MIN: gs_bucket<T, NEKOAMD_GS_MIN>(u, g, b, skip); break;
      case NEKOAMD_GS_MAX: gs_bucket<T, NEKOAMD_GS_MAX>(u, g, b, skip); break;
      default: return fail_arg("gs_op", "unknown gs operation");
    }
    NK_LAUNCHED();
  }
  return kOk;
}

template int gs_launch<double>(void*, double*, int, const int*);
template int gs_launch<float>(void*, float*, int, const int*);

template <typename T>
int gs_launch3(void* gs, T* const* u, int op, const int* const* skip,
               const bool* active) {
  const GsMap* g = (const GsMap*)gs;
  if (!g) return fail_arg("gs_op", "null gs handle");
  if (g->nblk == 0 || op != NEKOAMD_GS_ADD) {
    for (int c = 0; c < 3; ++c) {
      if (!active[c]) continue;
      const int rc = gs_launch<T>(gs, u[c], op, skip[c]);
      if (rc) return rc;
    }
    return kOk;
  }
  Gs3Vec<T> v{};
  for (int c = 0; c < 3; ++c) {
    if (active[c] && !u[c] && g->n > 0) return fail_arg("gs_op", "null field");
    v.u[c] = u[c];
    v.skip[c] = skip[c];
    v.active[c] = active[c] ? 1 : 0;
  }
  nk_launch(gs_blk3_k<T, NEKOAMD_GS_ADD>, dim3(g->nblk), dim3(256), 0, v,
            g->blk_idx, g->blk_desc);
  NK_LAUNCHED();
  for (int b = 0; b < g->nb; ++b) {
    if (g->cnt[b] == 0 || g->m[b] == 2 || g->m[b] == 4 || g->m[b] == 8)
      continue;
    for (int c = 0; c < 3; ++c) {
      if (!active[c]) continue;
      gs_bucket<T, NEKOAMD_GS_ADD>(u[c], g, b, skip[c]);
      NK_LAUNCHED();
    }
  }
  return kOk;
}

template int gs_launch3<double>(void*, double* const*, int, const int* const*,
                                const bool*);
template int gs_launch3<float>(void*, float* const*, int, const int* const*,
                               const bool*);

// Local points per group of the launch layout, and CUs per XCD on gfx950
// (MI355X: 256 CUs in 8 XCDs).
constexpr int64_t kBlkGroup = 8192;
constexpr int kXcdCus = 32;

static int xcd_count() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                            dev) != hipSuccess)
    return 1;
  return cus >= 2 * kXcdCus && cus % kXcdCus == 0 ? cus / kXcdCus : 1;
}

// The launch layout (gs_blk_k) of the canonical map's 2-, 4- and 8-buckets;
// `idx` is the canonical map (buckets at host_off).
static int build_blocked(GsMap* g, const std::vector<int32_t>& idx) {
  struct Blk {
    int m;
    int b;          // bucket
    int64_t first;  // position of its first tuple in `order`
    int cnt;
  };
  // each local point is the first (smallest) member of at most one tuple:
  // walking the points in order lists the tuples by first member, so the
  // lanes of a wave take neighbouring rows of the same elements
  std::vector<int32_t> at(g->n, -1);
  std::vector<int8_t> at_b(g->n, -1);
  for (int b = 0; b < g->nb; ++b) {
    const int m = g->m[b];
    if (m != 2 && m != 4 && m != 8) continue;
    const int32_t* q = idx.data() + g->host_off[b];
    for (int64_t t = 0; t < g->cnt[b]; ++t) {
      at[q[t * m]] = (int32_t)t;
      at_b[q[t * m]] = (int8_t)b;
    }
  }
  std::vector<int64_t> order;  // (bucket, tuple) of each block, in order
  std::vector<Blk> blocks;
  std::vector<int64_t> list[3];
  for (int64_t k0 = 0; k0 < g->n; k0 += kBlkGroup) {
    const int64_t k1 = std::min<int64_t>(g->n, k0 + kBlkGroup);
    for (auto& v : list) v.clear();
    for (int64_t l = k0; l < k1; ++l) {
      if (at[l] < 0) continue;
      const int m = g->m[at_b[l]];
      list[m == 2 ? 0 : m == 4 ? 1 : 2].push_back((int64_t)at_b[l] << 32 | at[l]);
    }
    for (int c = 0; c < 3; ++c)
      for (size_t t = 0; t < list[c].size(); t += 256) {
        const int cnt = (int)std::min<size_t>(256, list[c].size() - t);
        blocks.push_back({2 << c, (int)(list[c][t] >> 32), (int64_t)order.size(), cnt});
        order.insert(order.end(), list[c].begin() + t, list[c].begin() + t + cnt);
      }
  }
  if (blocks.empty()) return kOk;
  const int nxcd = xcd_count();
  const int64_t nlog = (int64_t)blocks.size();
  const int64_t per = (nlog + nxcd - 1) / nxcd;
  const int64_t ntab = per * nxcd;
  if (ntab > INT32_MAX) return fail_arg("gs_init", "launch layout too large");
  // tuple indices block by block in logical order, each block 16-byte aligned
  std::vector<int64_t> boff(nlog);
  int64_t len = 0;
  for (int64_t L = 0; L < nlog; ++L) {
    boff[L] = len;
    len += ((int64_t)blocks[L].cnt * blocks[L].m + 3) / 4 * 4;
  }
  if (len / 4 > INT32_MAX) return fail_arg("gs_init", "launch layout too large");
  std::vector<int32_t> bidx(len, 0);
  for (int64_t L = 0; L < nlog; ++L) {
    const Blk& k = blocks[L];
    const int32_t* q = idx.data() + g->host_off[k.b];
    for (int c = 0; c < k.cnt; ++c) {
      const int64_t t = order[k.first + c] & 0xffffffff;
      std::memcpy(bidx.data() + boff[L] + (int64_t)c * k.m, q + t * k.m,
                  sizeof(int32_t) * k.m);
    }
  }
  // workgroup r * nxcd + x (on XCD x) takes logical block x * per + r
  std::vector<int4> desc(ntab, int4{0, 0, 0, 0});
  int64_t tuples = 0;
  for (int64_t L = 0; L < nlog; ++L) {
    const int64_t x = L / per, r = L % per;
    if (tuples > INT32_MAX) return fail_arg("gs_init", "launch layout too large");
    desc[r * nxcd + x] =
        int4{blocks[L].m, (int)(boff[L] / 4), blocks[L].cnt, (int)tuples};
    tuples += blocks[L].cnt;
  }
  NK_HIP(hipMalloc((void**)&g->blk_idx, sizeof(int32_t) * len));
  NK_HIP(hipMalloc((void**)&g->blk_desc, sizeof(int4) * ntab));
  NK_HIP(hipMemcpy(g->blk_idx, bidx.data(), sizeof(int32_t) * len,
                   hipMemcpyHostToDevice));
  NK_HIP(hipMemcpy(g->blk_desc, desc.data(), sizeof(int4) * ntab,
                   hipMemcpyHostToDevice));
  g->nblk = (int)ntab;
  g->blk_len = len;
  g->blk_tuples = tuples;
  return kOk;
}

static void free_map(GsMap* g) {
  if (g->idx_d) (void)hipFree(g->idx_d);
  if (g->blk_idx) (void)hipFree(g->blk_idx);
  if (g->blk_desc) (void)hipFree(g->blk_desc);
  if (g->avg_b) (void)hipFree(g->avg_b);
  if (g->avg_binv) (void)hipFree(g->avg_binv);
  if (g->avg_bb) (void)hipFree(g->avg_bb);
  if (g->avg_binvb) (void)hipFree(g->avg_binvb);
  delete g;
}

}  // namespace nekoamd

using namespace nekoamd;

extern "C" {

int nekoamd_gs_init(void** gs, const int64_t* gid, const int64_t* n) {
  NK_INIT();
  NK_REQUIRE(gs && n && *n >= 0 && (gid || *n == 0), "bad gs_init arguments");
  std::vector<int> bm;
  std::vector<int64_t> bcnt;
  std::vector<int32_t> idx;
  int rc = build_gs_map(*n, gid, bm, bcnt, idx);
  if (rc != kOk) return rc;
  GsMap* g = new GsMap();
  g->n = *n;
  g->nb = (int)bm.size();
  int64_t hoff = 0, doff = 0;
  for (int b = 0; b < g->nb; ++b) {
    g->m[b] = bm[b];
    g->cnt[b] = bcnt[b];
    g->host_off[b] = hoff;
    g->dev_off[b] = doff;
    hoff += bcnt[b] * bm[b];
    doff += (bcnt[b] * bm[b] + 3) / 4 * 4;  // next bucket 16-byte aligned
  }
  g->nocc = hoff;
  g->dev_len = doff;
  if (doff > 0) {
    std::vector<int32_t> dev(doff, 0);
    for (int b = 0; b < g->nb; ++b)
      std::memcpy(dev.data() + g->dev_off[b], idx.data() + g->host_off[b],
                  sizeof(int32_t) * g->cnt[b] * g->m[b]);
    hipError_t e = hipMalloc((void**)&g->idx_d, sizeof(int32_t) * doff);
    if (e != hipSuccess) {
      delete g;
      return fail_hip(__func__, e);
    }
    e = hipMemcpy(g->idx_d, dev.data(), sizeof(int32_t) * doff,
                  hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      free_map(g);
      return fail_hip(__func__, e);
    }
    rc = build_blocked(g, idx);
    if (rc != kOk) {
      free_map(g);
      return rc;
    }
  }
  *gs = g;
  return kOk;
}

int nekoamd_gs_free(void* gs) {
  GsMap* g = (GsMap*)gs;
  if (!g) return kOk;
  NK_HIP(hipStreamSynchronize(stream()));
  free_map(g);
  return kOk;
}

int nekoamd_gs_op_f64(void* gs, void* u, const int* op) {
  NK_INIT();
  NK_REQUIRE(op, "null op");
  return gs_launch<double>(gs, (double*)u, *op, nullptr);
}

int nekoamd_gs_op_f32(void* gs, void* u, const int* op) {
  NK_INIT();
  NK_REQUIRE(op, "null op");
  return gs_launch<float>(gs, (float*)u, *op, nullptr);
}

int nekoamd_gs_mass_average(void* gs, void* u, const void* B,
                            const void* Binv) {
  NK_INIT();
  GsMap* g = (GsMap*)gs;
  NK_REQUIRE(g, "null gs handle");
  NK_REQUIRE(g->n == 0 || (u && B && Binv), "null field");
  if (g->dev_len == 0) return kOk;
  if (g->avg_key[0] != B || g->avg_key[1] != Binv) {
    int64_t groups = 0;
    for (int b = 0; b < g->nb; ++b) {
      g->goff[b] = groups;
      groups += g->cnt[b];
    }
    if (!g->avg_b) {
      NK_HIP(hipMalloc((void**)&g->avg_b, sizeof(double) * g->dev_len));
      NK_HIP(hipMalloc((void**)&g->avg_binv, sizeof(double) * (groups > 0 ? groups : 1)));
    }
    hipLaunchKernelGGL(gs_avg_gather_b_k, dim3(nk_grid(g->dev_len, 256)),
                       dim3(256), 0, stream(), g->dev_len, g->idx_d,
                       (const double*)B, g->avg_b);
    NK_LAUNCHED();
    for (int b = 0; b < g->nb; ++b) {
      if (g->cnt[b] == 0) continue;
      hipLaunchKernelGGL(gs_avg_gather_binv_k, dim3(nk_grid(g->cnt[b], 256)),
                         dim3(256), 0, stream(), g->cnt[b], g->m[b],
                         g->idx_d + g->dev_off[b], (const double*)Binv,
                         g->avg_binv + g->goff[b]);
      NK_LAUNCHED();
    }
    if (g->nblk > 0) {
      if (!g->avg_bb) {
        NK_HIP(hipMalloc((void**)&g->avg_bb, sizeof(double) * g->blk_len));
        NK_HIP(hipMalloc((void**)&g->avg_binvb, sizeof(double) * g->blk_tuples));
      }
      hipLaunchKernelGGL(gs_avg_gather_b_k, dim3(nk_grid(g->blk_len, 256)),
                         dim3(256), 0, stream(), g->blk_len, g->blk_idx,
                         (const double*)B, g->avg_bb);
      NK_LAUNCHED();
      hipLaunchKernelGGL(gs_avg_gather_binv_blk_k, dim3(g->nblk), dim3(256),
                         0, stream(), g->blk_idx, g->blk_desc,
                         (const double*)Binv, g->avg_binvb);
      NK_LAUNCHED();
    }
    g->avg_key[0] = B;
    g->avg_key[1] = Binv;
  }
  AvgBucket bk[3] = {};
  for (int b = 0; b < g->nb; ++b) {
    const int slot = g->m[b] == 2 ? 0 : g->m[b] == 4 ? 1 : g->m[b] == 8 ? 2 : -1;
    const AvgBucket ab{g->idx_d + g->dev_off[b], g->avg_b + g->dev_off[b],
                       g->avg_binv + g->goff[b], g->cnt[b]};
    if (slot >= 0) {
      bk[slot] = ab;
    } else if (g->cnt[b] > 0) {
      hipLaunchKernelGGL(gs_avg_any_k, dim3((unsigned)((g->cnt[b] + 255) / 256)),
                         dim3(256), 0, stream(), (double*)u, ab.bq, ab.binv,
                         ab.idx, ab.cnt, g->m[b]);
      NK_LAUNCHED();
    }
  }
  const int64_t tot = bk[0].cnt + bk[1].cnt + bk[2].cnt;
  if (g->nblk > 0) {
    hipLaunchKernelGGL(gs_avg_blk_k, dim3(g->nblk), dim3(256), 0, stream(),
                       (double*)u, g->blk_idx, g->avg_bb, g->avg_binvb,
                       g->blk_desc);
    NK_LAUNCHED();
  } else if (tot > 0) {
    hipLaunchKernelGGL(gs_avg_248_k, dim3((unsigned)((tot + 255) / 256)),
                       dim3(256), 0, stream(), (double*)u, bk[0], bk[1], bk[2]);
    NK_LAUNCHED();
  }
  return kOk;
}

int nekoamd_gs_map_info(void* gs, int* nb, int* bm, int64_t* bcnt,
                        int64_t* nocc) {
  const GsMap* g = (const GsMap*)gs;
  NK_REQUIRE(g && nb && nocc, "null argument");
  *nb = g->nb;
  *nocc = g->nocc;
  for (int b = 0; b < g->nb; ++b) {
    if (bm) bm[b] = g->m[b];
    if (bcnt) bcnt[b] = g->cnt[b];
  }
  return kOk;
}

int nekoamd_gs_map_copy(void* gs, int32_t* idx) {
  NK_INIT();
  const GsMap* g = (const GsMap*)gs;
  NK_REQUIRE(g && (idx || g->nocc == 0), "null argument");
  for (int b = 0; b < g->nb; ++b)
    NK_HIP(hipMemcpy(idx + g->host_off[b], g->idx_d + g->dev_off[b],
                     sizeof(int32_t) * g->cnt[b] * g->m[b],
                     hipMemcpyDeviceToHost));
  return kOk;
}

}  // extern "C"
