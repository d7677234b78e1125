// math.hip -- FP64 vector kernels on the path (Neko device_math as the
// reference calls it: field_subcol3 in the Brinkman term,
// sources/source_terms/simple_brinkman_source_term.f90:149-151; col3 /
// cfill in the filter, sources/mapping_functions/PDE_filter_mapping.f90:225-241;
// glsc2 in the objective and volume, sources/objectives/*.f90), the
// Dirichlet bc zeroing, and the reference's own masked kernels
// (sources/neko_ext/math/bcknd/device/cuda/math_ext_kernel.h:42-114,
// 1-based Fortran masks) plus glsc2_mask (math_ext.f90:99-118).
//
// All are grid-stride streams; grids are sized to cover the 256 CUs many
// times (nk_grid) and the reductions are deterministic (reduce.h).
#include "internal.h"
#include "nekoamd.h"
#include "reduce.h"

namespace nekoamd {

// Reduction scratch: kRedGrid partials + one result, device; one pinned
// host double for the read-back.
static double* g_part = nullptr;
static double* g_host_res = nullptr;

static int red_scratch() {
  if (!g_part) {
    NK_HIP(hipMalloc((void**)&g_part, sizeof(double) * (kRedGrid + 8)));
    NK_HIP(hipHostMalloc((void**)&g_host_res, sizeof(double) * 8,
                         hipHostMallocDefault));
  }
  return kOk;
}

constexpr int kBlock = 256;

template <typename F>
__global__ void __launch_bounds__(kBlock) map_k(int64_t n, F f) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    f(i);
}

template <typename F>
static int map_launch(int64_t n, F f) {
  if (n <= 0) return kOk;
  const int g = nk_grid(n, kBlock, kMapGridCap);
  hipLaunchKernelGGL((map_k<F>), dim3(g), dim3(kBlock), 0,
                     stream(), n, f);
  NK_LAUNCHED();
  return kOk;
}

// Partial sums of f(i) over a grid-stride range: one partial per block.
template <typename F>
__global__ void __launch_bounds__(kRedBlock)
    sum_k(int64_t n, F f, double* __restrict__ part) {
  __shared__ double sh[kRedBlock / 64];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kRedBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRedBlock)
    s += f(i);
  s = block_sum<kRedBlock>(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kFinBlock)
    finalize_k(const double* __restrict__ part, int np, double* out) {
  __shared__ double sh[kFinBlock / 64];
  const double s = sum_partials<kFinBlock>(part, np, sh);
  if (threadIdx.x == 0) *out = s;
}

template <typename F>
static int sum_launch(int64_t n, F f, double* res) {
  int rc = red_scratch();
  if (rc) return rc;
  if (n <= 0) {
    *res = 0.0;
    return kOk;
  }
  const int g = nk_grid(n, kRedBlock, kRedGrid);
  hipLaunchKernelGGL((sum_k<F>), dim3(g), dim3(kRedBlock), 0, stream(), n, f,
                     g_part);
  NK_LAUNCHED();
  hipLaunchKernelGGL(finalize_k, dim3(1), dim3(kFinBlock), 0, stream(),
                     g_part, g, g_part + kRedGrid);
  NK_LAUNCHED();
  if (const int rc = to_host(g_host_res, g_part + kRedGrid, sizeof(double)))
    return rc;
  NK_HIP(hipStreamSynchronize(stream()));
  *res = *g_host_res;
  return kOk;
}

template <typename T>
int bc_zero_launch(T* u, const int32_t* mask, int64_t nmask, const int* skip) {
  if (nmask <= 0) return kOk;
  auto f = [=] __device__(int64_t i) {
    if (skip != nullptr && *skip) return;
    u[mask[i]] = T(0);
  };
  return map_launch(nmask, f);
}

template int bc_zero_launch<double>(double*, const int32_t*, int64_t,
                                    const int*);
template int bc_zero_launch<float>(float*, const int32_t*, int64_t,
                                   const int*);

}  // namespace nekoamd

using namespace nekoamd;

#define NK_N(n) \
  NK_INIT();    \
  NK_REQUIRE(n && *n >= 0, "bad length")

extern "C" {

int nekoamd_copy(void* a, const void* b, const int64_t* n) {
  NK_N(n);
  NK_REQUIRE(*n == 0 || (a && b), "null field");
  NK_HIP(hipMemcpyAsync(a, b, sizeof(double) * (size_t)*n,
                        hipMemcpyDeviceToDevice, stream()));
  return kOk;
}

int nekoamd_rzero(void* a, const int64_t* n) {
  NK_N(n);
  double* A = (double*)a;
  return map_launch(*n, [=] __device__(int64_t i) { A[i] = 0.0; });
}

int nekoamd_cfill(void* a, const double* c, const int64_t* n) {
  NK_N(n);
  NK_REQUIRE(c, "null scalar");
  double* A = (double*)a;
  const double C = *c;
  return map_launch(*n, [=] __device__(int64_t i) { A[i] = C; });
}

int nekoamd_cmult(void* a, const double* c, const int64_t* n) {
  NK_N(n);
  NK_REQUIRE(c, "null scalar");
  double* A = (double*)a;
  const double C = *c;
  return map_launch(*n, [=] __device__(int64_t i) { A[i] = C * A[i]; });
}

int nekoamd_add2(void* a, const void* b, const int64_t* n) {
  NK_N(n);
  double* A = (double*)a;
  const double* Bv = (const double*)b;
  return map_launch(*n, [=] __device__(int64_t i) { A[i] = A[i] + Bv[i]; });
}

int nekoamd_cadd2(void* a, const void* b, const double* c, const int64_t* n) {
  NK_N(n);
  NK_REQUIRE(c, "null scalar");
  double* A = (double*)a;
  const double* Bv = (const double*)b;
  const double C = *c;
  return map_launch(*n, [=] __device__(int64_t i) { A[i] = Bv[i] + C; });
}

int nekoamd_add2s2(void* a, const void* b, const double* c, const int64_t* n) {
  NK_N(n);
  NK_REQUIRE(c, "null scalar");
  double* A = (double*)a;
  const double* Bv = (const double*)b;
  const double C = *c;
  return map_launch(*n,
                    [=] __device__(int64_t i) { A[i] = A[i] + C * Bv[i]; });
}

int nekoamd_col2(void* a, const void* b, const int64_t* n) {
  NK_N(n);
  double* A = (double*)a;
  const double* Bv = (const double*)b;
  return map_launch(*n, [=] __device__(int64_t i) { A[i] = A[i] * Bv[i]; });
}

int nekoamd_col3(void* a, const void* b, const void* c, const int64_t* n) {
  NK_N(n);
  double* A = (double*)a;
  const double *Bv = (const double*)b, *Cv = (const double*)c;
  return map_launch(*n, [=] __device__(int64_t i) { A[i] = Bv[i] * Cv[i]; });
}

int nekoamd_addcol3(void* a, const void* b, const void* c, const int64_t* n) {
  NK_N(n);
  double* A = (double*)a;
  const double *Bv = (const double*)b, *Cv = (const double*)c;
  return map_launch(*n,
                    [=] __device__(int64_t i) { A[i] = A[i] + Bv[i] * Cv[i]; });
}

int nekoamd_subcol3(void* a, const void* b, const void* c, const int64_t* n) {
  NK_N(n);
  double* A = (double*)a;
  const double *Bv = (const double*)b, *Cv = (const double*)c;
  return map_launch(*n,
                    [=] __device__(int64_t i) { A[i] = A[i] - Bv[i] * Cv[i]; });
}

int nekoamd_invcol1(void* a, const int64_t* n) {
  NK_N(n);
  double* A = (double*)a;
  return map_launch(*n, [=] __device__(int64_t i) { A[i] = 1.0 / A[i]; });
}

int nekoamd_f64_to_f32(void* a32, const void* b64, const int64_t* n) {
  NK_N(n);
  float* A = (float*)a32;
  const double* Bv = (const double*)b64;
  return map_launch(*n, [=] __device__(int64_t i) { A[i] = (float)Bv[i]; });
}

int nekoamd_f32_to_f64(void* a64, const void* b32, const int64_t* n) {
  NK_N(n);
  double* A = (double*)a64;
  const float* Bv = (const float*)b32;
  return map_launch(*n, [=] __device__(int64_t i) { A[i] = (double)Bv[i]; });
}

int nekoamd_bc_zero(void* u, const void* mask, const int64_t* nmask) {
  NK_N(nmask);
  NK_REQUIRE(*nmask == 0 || (u && mask), "null argument");
  return bc_zero_launch<double>((double*)u, (const int32_t*)mask, *nmask,
                                nullptr);
}

int nekoamd_bc_zero_f32(void* u, const void* mask, const int64_t* nmask) {
  NK_N(nmask);
  NK_REQUIRE(*nmask == 0 || (u && mask), "null argument");
  return bc_zero_launch<float>((float*)u, (const int32_t*)mask, *nmask,
                               nullptr);
}

int nekoamd_glsc2(double* res, const void* a, const void* b, const int64_t* n) {
  NK_N(n);
  NK_REQUIRE(res, "null result");
  const double *A = (const double*)a, *Bv = (const double*)b;
  return sum_launch(*n, [=] __device__(int64_t i) { return A[i] * Bv[i]; },
                    res);
}

int nekoamd_glsc3(double* res, const void* a, const void* b, const void* c,
                  const int64_t* n) {
  NK_N(n);
  NK_REQUIRE(res, "null result");
  const double *A = (const double*)a, *Bv = (const double*)b,
               *Cv = (const double*)c;
  return sum_launch(
      *n, [=] __device__(int64_t i) { return A[i] * Bv[i] * Cv[i]; }, res);
}

// ---- the reference's masked kernels (1-based masks) -----------------------

#define NK_MASK_ARGS(size, mask, mask_size)                           \
  NK_INIT();                                                          \
  NK_REQUIRE(size && mask_size && *mask_size >= 0, "bad mask size"); \
  NK_REQUIRE(*mask_size == 0 || mask, "null mask")

int hip_cadd_mask(void* a, const double* c, const int* size, const void* mask,
                  const int* mask_size) {
  NK_MASK_ARGS(size, mask, mask_size);
  NK_REQUIRE(c, "null scalar");
  double* A = (double*)a;
  const int* M = (const int*)mask;
  const double C = *c;
  return map_launch(*mask_size, [=] __device__(int64_t i) {
    const int m = M[i] - 1;
    A[m] = A[m] + C;
  });
}

int hip_invcol1_mask(void* a, const int* size, const void* mask,
                     const int* mask_size) {
  NK_MASK_ARGS(size, mask, mask_size);
  double* A = (double*)a;
  const int* M = (const int*)mask;
  return map_launch(*mask_size, [=] __device__(int64_t i) {
    const int m = M[i] - 1;
    A[m] = 1.0 / A[m];
  });
}

int hip_col2_mask(void* a, const void* b, const int* size, const void* mask,
                  const int* mask_size) {
  NK_MASK_ARGS(size, mask, mask_size);
  double* A = (double*)a;
  const double* Bv = (const double*)b;
  const int* M = (const int*)mask;
  return map_launch(*mask_size, [=] __device__(int64_t i) {
    const int m = M[i] - 1;
    A[m] = A[m] * Bv[m];
  });
}

int hip_col3_mask(void* a, const void* b, const void* c, const int* size,
                  const void* mask, const int* mask_size) {
  NK_MASK_ARGS(size, mask, mask_size);
  double* A = (double*)a;
  const double *Bv = (const double*)b, *Cv = (const double*)c;
  const int* M = (const int*)mask;
  return map_launch(*mask_size, [=] __device__(int64_t i) {
    const int m = M[i] - 1;
    A[m] = Bv[m] * Cv[m];
  });
}

int hip_sub3_mask(void* a, const void* b, const void* c, const int* size,
                  const void* mask, const int* mask_size) {
  NK_MASK_ARGS(size, mask, mask_size);
  double* A = (double*)a;
  const double *Bv = (const double*)b, *Cv = (const double*)c;
  const int* M = (const int*)mask;
  return map_launch(*mask_size, [=] __device__(int64_t i) {
    const int m = M[i] - 1;
    A[m] = Bv[m] - Cv[m];
  });
}

int hip_glsc2_mask(double* res, const void* a, const void* b, const int* size,
                   const void* mask, const int* mask_size) {
  NK_MASK_ARGS(size, mask, mask_size);
  NK_REQUIRE(res, "null result");
  const double *A = (const double*)a, *Bv = (const double*)b;
  const int* M = (const int*)mask;
  return sum_launch(
      *mask_size,
      [=] __device__(int64_t i) {
        const int m = M[i] - 1;
        return A[m] * Bv[m];
      },
      res);
}

}  // extern "C"
