// runtime.hip -- device set-up, memory and the library stream.
//
// Mirrors the device layer the reference reaches through Neko
// (device_map / device_memcpy / device_free, e.g.
// sources/topology_optimization/design.f90:150-155, :189-191, :214) and the
// single global in-order command queue its kernels run on
// (sources/neko_ext/math/bcknd/device/cuda/math_ext.cu:54, glb_cmd_queue).
#include <cstring>

#include "internal.h"
#include "nekoamd.h"

namespace nekoamd {

static thread_local std::string g_err;
static hipStream_t g_stream = nullptr;
static int g_device = -1;

void set_error(const char* where, const char* what) {
  g_err = std::string(where) + ": " + what;
}

int fail_hip(const char* where, hipError_t e) {
  set_error(where, hipGetErrorString(e));
  return kErrHip;
}

int fail_arg(const char* where, const char* what) {
  set_error(where, what);
  return kErrArg;
}

hipStream_t stream() { return g_stream; }
LaunchRecorder*& launch_recorder() {
  static thread_local LaunchRecorder* rec = nullptr;
  return rec;
}
hipStream_t& launch_stream() {
  static thread_local hipStream_t st = nullptr;
  return st;
}
bool initialised() { return g_stream != nullptr; }
int current_device() { return g_device; }

// Each piece copied word by word by one workgroup; the pinned host pages are
// written directly from the device.
__global__ void __launch_bounds__(256) to_host_k(HostCopy c) {
  for (int q = 0; q < c.n; ++q) {
    auto* d = (uint32_t*)c.dst[q];
    const auto* s = (const uint32_t*)c.src[q];
    for (int w = threadIdx.x; w < c.words[q]; w += 256) d[w] = s[w];
  }
}

int to_host(const HostCopy& c) {
  for (int q = 0; q < c.n; ++q)
    if (c.words[q] < 0 || ((uintptr_t)c.dst[q] | (uintptr_t)c.src[q]) % 4 != 0)
      return fail_arg("to_host", "pieces must be 4-byte aligned");
  hipLaunchKernelGGL(to_host_k, dim3(1), dim3(256), 0, stream(), c);
  NK_LAUNCHED();
  return kOk;
}

// cnt = 1 / mult as a byte where mult is exactly 1.0 / c for an integer c in
// [1, 255]; every other point adds one to *bad.
__global__ void __launch_bounds__(256)
    mult_count_k(int64_t n, const double* __restrict__ mult,
                 uint8_t* __restrict__ cnt, int* __restrict__ bad) {
  for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * 256) {
    const double m = mult[p];
    const double c = rint(1.0 / m);
    const bool ok = c >= 1.0 && c <= 255.0 && 1.0 / c == m;
    cnt[p] = ok ? (uint8_t)c : (uint8_t)1;
    if (!ok) atomicAdd(bad, 1);
  }
}

__global__ void zero_int_k(int* __restrict__ v) { *v = 0; }

int mult_counts(int64_t n, const double* mult, uint8_t* cnt, int* bad,
                int* host, bool* ok) {
  hipLaunchKernelGGL(zero_int_k, dim3(1), dim3(1), 0, stream(), bad);
  hipLaunchKernelGGL(mult_count_k, dim3(nk_grid(n, 256)), dim3(256), 0,
                     stream(), n, mult, cnt, bad);
  NK_LAUNCHED();
  if (const int rc = to_host(host, bad, sizeof(int))) return rc;
  NK_HIP(hipStreamSynchronize(stream()));
  *ok = host[0] == 0;
  return kOk;
}

int to_host(void* dst, const void* src, size_t bytes) {
  if (bytes % 4 != 0) return fail_arg("to_host", "size is not a multiple of 4");
  HostCopy c{};
  c.n = 1;
  c.dst[0] = dst;
  c.src[0] = src;
  c.words[0] = (int)(bytes / 4);
  return to_host(c);
}

}  // namespace nekoamd

using namespace nekoamd;

extern "C" {

const char* nekoamd_error_string(void) { return g_err.c_str(); }

int nekoamd_device_init(const int* device) {
  NK_REQUIRE(device, "device is null");
  int ndev = 0;
  NK_HIP(hipGetDeviceCount(&ndev));
  NK_REQUIRE(*device >= 0 && *device < ndev, "no such device");
  if (g_stream && g_device == *device) return kOk;
  NK_HIP(hipSetDevice(*device));
  const int rc = objective_scratch_setup(*device);
  if (rc) return rc;
  if (g_stream) NK_HIP(hipStreamDestroy(g_stream));
  // Default priority, non-blocking with respect to the null stream.
  NK_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  g_device = *device;
  return kOk;
}

int nekoamd_device_finalize(void) {
  if (!g_stream) return kOk;
  NK_HIP(hipStreamSynchronize(g_stream));
  NK_HIP(hipStreamDestroy(g_stream));
  g_stream = nullptr;
  g_device = -1;
  return kOk;
}

int nekoamd_device_name(char* buf, const int* len) {
  NK_REQUIRE(buf && len && *len > 0, "bad buffer");
  hipDeviceProp_t prop;
  int dev = g_device < 0 ? 0 : g_device;
  NK_HIP(hipGetDeviceProperties(&prop, dev));
  std::snprintf(buf, (size_t)*len, "%s (%s, %d CUs)", prop.name,
                prop.gcnArchName, prop.multiProcessorCount);
  return kOk;
}

void* nekoamd_stream(void) { return (void*)g_stream; }

int nekoamd_device_alloc(void** ptr, const int64_t* bytes) {
  NK_INIT();
  NK_REQUIRE(ptr && bytes && *bytes >= 0, "bad alloc arguments");
  *ptr = nullptr;
  if (*bytes == 0) return kOk;
  NK_HIP(hipMalloc(ptr, (size_t)*bytes));
  return kOk;
}

int nekoamd_device_free(void* ptr) {
  if (!ptr) return kOk;
  NK_HIP(hipStreamSynchronize(g_stream));
  NK_HIP(hipFree(ptr));
  return kOk;
}

int nekoamd_host_alloc_pinned(void** ptr, const int64_t* bytes) {
  NK_REQUIRE(ptr && bytes && *bytes > 0, "bad alloc arguments");
  NK_HIP(hipHostMalloc(ptr, (size_t)*bytes, hipHostMallocDefault));
  return kOk;
}

int nekoamd_host_free_pinned(void* ptr) {
  if (!ptr) return kOk;
  NK_HIP(hipHostFree(ptr));
  return kOk;
}

// kind: NEKOAMD_HOST_TO_DEVICE, NEKOAMD_DEVICE_TO_HOST, NEKOAMD_DEVICE_TO_DEVICE
int nekoamd_device_memcpy(void* dst, const void* src, const int64_t* bytes,
                          const int* kind, const int* sync) {
  NK_INIT();
  NK_REQUIRE(dst && src && bytes && kind && sync, "null argument");
  if (*bytes == 0) return kOk;
  hipMemcpyKind k;
  switch (*kind) {
    case NEKOAMD_HOST_TO_DEVICE: k = hipMemcpyHostToDevice; break;
    case NEKOAMD_DEVICE_TO_HOST: k = hipMemcpyDeviceToHost; break;
    case NEKOAMD_DEVICE_TO_DEVICE: k = hipMemcpyDeviceToDevice; break;
    default: return fail_arg(__func__, "unknown copy kind");
  }
  NK_HIP(hipMemcpyAsync(dst, src, (size_t)*bytes, k, g_stream));
  if (*sync) NK_HIP(hipStreamSynchronize(g_stream));
  return kOk;
}

int nekoamd_device_sync(void) {
  NK_INIT();
  NK_HIP(hipStreamSynchronize(g_stream));
  return kOk;
}

int nekoamd_event_create(void** ev) {
  NK_REQUIRE(ev, "null event");
  hipEvent_t e;
  NK_HIP(hipEventCreate(&e));
  *ev = (void*)e;
  return kOk;
}

int nekoamd_event_destroy(void* ev) {
  if (ev) NK_HIP(hipEventDestroy((hipEvent_t)ev));
  return kOk;
}

int nekoamd_event_record(void* ev) {
  NK_INIT();
  NK_HIP(hipEventRecord((hipEvent_t)ev, g_stream));
  return kOk;
}

int nekoamd_event_sync(void* ev) {
  NK_HIP(hipEventSynchronize((hipEvent_t)ev));
  return kOk;
}

int nekoamd_event_elapsed(void* ev0, void* ev1, float* ms) {
  NK_REQUIRE(ev0 && ev1 && ms, "null argument");
  NK_HIP(hipEventElapsedTime(ms, (hipEvent_t)ev0, (hipEvent_t)ev1));
  return kOk;
}

}  // extern "C"
