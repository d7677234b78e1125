// internal.h -- shared plumbing of libnekoamd: error reporting, the single
// in-order stream every kernel is queued on, launch checks.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <tuple>

namespace nekoamd {

// Status codes returned by every extern "C" entry point.
enum Status : int {
  kOk = 0,
  kErrHip = 1,       // a HIP runtime call failed
  kErrArg = 2,       // invalid argument (null pointer, size out of range)
  kErrState = 3,     // library not initialised / handle invalid
  kErrUnsupported = 4,
  kErrNumeric = 5,   // e.g. singular system
};

void set_error(const char* where, const char* what);
int fail_hip(const char* where, hipError_t e);
int fail_arg(const char* where, const char* what);
hipStream_t stream();  // the library stream (created by nekoamd_device_init)
bool initialised();
int current_device();

// Graph building without stream capture: while a LaunchRecorder is installed
// on the calling thread, nk_launch appends each launch as a kernel node of
// `graph` that depends on the node before it (an in-order chain, like the
// stream); otherwise nk_launch queues the kernel on the library stream (or
// on launch_stream() when one is set).  Only
// code whose every device operation goes through nk_launch may run under a
// recorder.
struct LaunchRecorder {
  hipGraph_t graph = nullptr;
  hipGraphNode_t last = nullptr;
  hipError_t err = hipSuccess;
  int nodes = 0;
};
LaunchRecorder*& launch_recorder();  // per thread, null when not recording
// Per thread: the stream nk_launch queues on when not recording (null: the
// library stream).  Set around a sequence of launches that belongs on a side
// stream, and reset after it.
hipStream_t& launch_stream();

// Small device-to-host copies into pinned host memory (hipHostMalloc) by a
// kernel on the library stream instead of copy commands: up to 4 pieces
// (4-byte aligned, sizes in 4-byte words) in one launch.  The host reads
// them once the stream or an event recorded after them has completed.
struct HostCopy {
  void* dst[4];
  const void* src[4];
  int words[4];
  int n;
};
int to_host(const HostCopy& c);
int to_host(void* dst, const void* src, size_t bytes);

// The multiplicity weights of a mesh as one-byte counts: cnt[p] = c where
// mult[p] is exactly 1.0 / c for an integer c in [1, 255]; *ok when every
// point is (then 1.0 / cnt[p] is the same double as mult[p]).  bad: one
// device int; host: pinned host memory for one int.  Synchronises.
int mult_counts(int64_t n, const double* mult, uint8_t* cnt, int* bad,
                int* host, bool* ok);

template <typename... P, typename... A>
inline void nk_launch(void (*kernel)(P...), dim3 grid, dim3 block,
                      unsigned shm, A... a) {
  static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
  // the arguments converted to the kernel's parameter types, and their
  // addresses in parameter order
  std::tuple<P...> args{static_cast<P>(a)...};
  void* ptrs[sizeof...(P) + 1];
  std::apply([&ptrs](auto&... v) {
    int i = 0;
    ((ptrs[i++] = (void*)&v), ...);
  }, args);
  LaunchRecorder* rec = launch_recorder();
  if (!rec) {
    // the error, if any, stays the thread's last error for NK_LAUNCHED
    hipStream_t st = launch_stream();
    (void)hipLaunchKernel((const void*)kernel, grid, block, ptrs, shm,
                          st ? st : stream());
    return;
  }
  if (rec->err != hipSuccess) return;
  hipKernelNodeParams np{};
  np.func = (void*)kernel;
  np.gridDim = grid;
  np.blockDim = block;
  np.sharedMemBytes = shm;
  np.kernelParams = ptrs;  // copied into the node by hipGraphAddKernelNode
  np.extra = nullptr;
  hipGraphNode_t node = nullptr;
  rec->err = hipGraphAddKernelNode(&node, rec->graph, rec->last ? &rec->last : nullptr,
                                   rec->last ? 1 : 0, &np);
  if (rec->err == hipSuccess) {
    rec->last = node;
    ++rec->nodes;
  }
}

}  // namespace nekoamd

#define NK_HIP(call)                                                    \
  do {                                                                  \
    hipError_t nk_e_ = (call);                                          \
    if (nk_e_ != hipSuccess) return nekoamd::fail_hip(__func__, nk_e_); \
  } while (0)

#define NK_LAUNCHED()                                                   \
  do {                                                                  \
    hipError_t nk_e_ = hipGetLastError();                               \
    if (nk_e_ != hipSuccess) return nekoamd::fail_hip(__func__, nk_e_); \
  } while (0)

#define NK_REQUIRE(cond, msg)                                      \
  do {                                                             \
    if (!(cond)) return nekoamd::fail_arg(__func__, msg);          \
  } while (0)

#define NK_INIT()                                                        \
  do {                                                                   \
    if (!nekoamd::initialised()) {                                       \
      nekoamd::set_error(__func__, "nekoamd_device_init was not called"); \
      return nekoamd::kErrState;                                         \
    }                                                                    \
  } while (0)

// Grid cap of the elementwise map kernels (map_k, pmap_k, tmap_k): one point
// per thread up to 2^28 points.  Measured on MI355X for an update of 5
// input and 3 output streams at 16.8 M points: 187 us with one point per
// thread, 215 us on 8192 blocks striding over the range.
constexpr int kMapGridCap = 1 << 20;

// Grid size for grid-stride elementwise kernels: enough waves to cover the
// 256 CUs several times over, never more blocks than work.
static inline int nk_grid(int64_t n, int block, int cap = 8192) {
  int64_t g = (n + block - 1) / block;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return (int)g;
}

namespace nekoamd {

// ax_helm launcher (plain entry points and the FP64 residual of the mixed
// CG).  skip: device flag that turns the launch into a no-op when non-zero
// (may be null).
template <typename T, typename TG>
int ax_helm_launch(T* w, const T* u, const double* dx, const T* h1,
                   const T* h2, const TG* g11, const TG* g12, const TG* g13,
                   const TG* g22, const TG* g23, const TG* g33, const TG* B,
                   int lx, int nelv, bool ifh2, const int* skip);
// w[c] = A u[c] for the three fields in one launch (FP64); field c is left
// alone if !active[c] or while *skip[c] is set (skip and active may be null:
// none skipped, all active).  With F (and G): w[c] = (F[c] - G[c]) - A u[c].
int ax_helm3_launch(double* const* w, const double* const* u,
                    const int* const* skip, const bool* active,
                    const double* dx, const double* h1, const double* h2,
                    const double* g11, const double* g12, const double* g13,
                    const double* g22, const double* g23, const double* g33,
                    const double* B, int lx, int nelv, bool ifh2,
                    const double* const* F = nullptr,
                    const double* const* G = nullptr);

// Gather-scatter on a map built by nekoamd_gs_init; skip as above.
template <typename T>
int gs_launch(void* gs, T* u, int op, const int* skip);
// gs_launch on three fields of the same map at once (the index data read
// once): field c is left alone if !active[c] or while *skip[c] is set.
template <typename T>
int gs_launch3(void* gs, T* const* u, int op, const int* const* skip,
               const bool* active);

template <typename T>
int bc_zero_launch(T* u, const int32_t* mask, int64_t nmask, const int* skip);

// Device address of the allocation-free objective's static scratch for
// `device` (topopt.hip); called by nekoamd_device_init with that device
// current.
int objective_scratch_setup(int device);

// The element network's continuous, masked correction direction for the
// residual r (mlp.hip); used inside the flexible CG (fcg.hip).
int mlp_direction(void* mlp, int64_t n, const double* r, double* q,
                  const double* mult, void* gs, const int32_t* mask,
                  int64_t nmask);

}  // namespace nekoamd
