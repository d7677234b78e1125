// chkp.cpp -- checkpoint files of the Pn/Pn time stepper: the state a run
// restarts from, as the reference's chkp_t keeps it (the fluid and its
// lagged fields, sources/adjoint/adjoint_scheme.f90:159; restored by
// adjoint_pnpn_restart, sources/adjoint/adjoint_pnpn.f90:418-546).  Neko's
// own chkp layout is Neko's and no file of it is in the reference, so the
// format is this library's:
//
//   char[8]   "NKAMDCHK"
//   int32     version (1)
//   int32     nfields
//   int64     n (values per field)
//   int32     tstep
//   int32     0
//   float64   time, dt
//   nfields x n float64 (little endian, the fields in the caller's order)
//
// Host code: device fields are copied through one pinned staging buffer of
// at most 64 MiB.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>

#include "internal.h"
#include "nekoamd.h"

namespace nekoamd {
namespace {

constexpr char kMagic[8] = {'N', 'K', 'A', 'M', 'D', 'C', 'H', 'K'};
constexpr int32_t kVersion = 1;
constexpr int64_t kStage = 8 << 20;  // values per staging chunk (64 MiB)

struct Header {
  char magic[8];
  int32_t version;
  int32_t nfields;
  int64_t n;
  int32_t tstep;
  int32_t zero;
  double time;
  double dt;
};
static_assert(sizeof(Header) == 48, "checkpoint header layout");

struct File {
  FILE* f = nullptr;
  ~File() {
    if (f) std::fclose(f);
  }
};

struct Pinned {
  double* p = nullptr;
  ~Pinned() {
    if (p) (void)hipHostFree(p);
  }
};

}  // namespace
}  // namespace nekoamd

using namespace nekoamd;

extern "C" {

int nekoamd_chkp_write(const char* path, const int* tstep, const double* time,
                       const double* dt, const int* nfields,
                       const void* const* fields, const int64_t* n) {
  NK_INIT();
  NK_REQUIRE(path && tstep && time && dt && nfields && fields && n,
             "null argument");
  NK_REQUIRE(*nfields > 0 && *n >= 0, "bad field count or size");
  for (int k = 0; k < *nfields; ++k)
    NK_REQUIRE(*n == 0 || fields[k], "null field");
  Header h{};
  std::memcpy(h.magic, kMagic, sizeof(kMagic));
  h.version = kVersion;
  h.nfields = *nfields;
  h.n = *n;
  h.tstep = *tstep;
  h.time = *time;
  h.dt = *dt;
  File file;
  file.f = std::fopen(path, "wb");
  if (!file.f) return fail_arg(__func__, "cannot open the checkpoint file for writing");
  if (std::fwrite(&h, sizeof(h), 1, file.f) != 1)
    return fail_arg(__func__, "write failed");
  Pinned buf;
  const int64_t chunk = *n < kStage ? *n : kStage;
  if (chunk > 0)
    NK_HIP(hipHostMalloc((void**)&buf.p, sizeof(double) * chunk,
                         hipHostMallocDefault));
  for (int k = 0; k < *nfields; ++k) {
    const double* src = (const double*)fields[k];
    for (int64_t o = 0; o < *n; o += chunk) {
      const int64_t m = *n - o < chunk ? *n - o : chunk;
      NK_HIP(hipMemcpyAsync(buf.p, src + o, sizeof(double) * m,
                            hipMemcpyDeviceToHost, stream()));
      NK_HIP(hipStreamSynchronize(stream()));
      if (std::fwrite(buf.p, sizeof(double), (size_t)m, file.f) != (size_t)m)
        return fail_arg(__func__, "write failed");
    }
  }
  if (std::fclose(file.f) != 0) {
    file.f = nullptr;
    return fail_arg(__func__, "close failed");
  }
  file.f = nullptr;
  return kOk;
}

// Reads a file written by nekoamd_chkp_write into nfields device fields of
// n values; the header must match nfields and n.
int nekoamd_chkp_read(const char* path, int* tstep, double* time, double* dt,
                      const int* nfields, void* const* fields,
                      const int64_t* n) {
  NK_INIT();
  NK_REQUIRE(path && tstep && time && dt && nfields && fields && n,
             "null argument");
  NK_REQUIRE(*nfields > 0 && *n >= 0, "bad field count or size");
  for (int k = 0; k < *nfields; ++k)
    NK_REQUIRE(*n == 0 || fields[k], "null field");
  File file;
  file.f = std::fopen(path, "rb");
  if (!file.f) return fail_arg(__func__, "cannot open the checkpoint file");
  Header h{};
  if (std::fread(&h, sizeof(h), 1, file.f) != 1)
    return fail_arg(__func__, "short checkpoint header");
  NK_REQUIRE(std::memcmp(h.magic, kMagic, sizeof(kMagic)) == 0,
             "not a checkpoint file");
  NK_REQUIRE(h.version == kVersion, "unsupported checkpoint version");
  NK_REQUIRE(h.nfields == *nfields && h.n == *n,
             "checkpoint holds a different number or size of fields");
  Pinned buf;
  const int64_t chunk = *n < kStage ? *n : kStage;
  if (chunk > 0)
    NK_HIP(hipHostMalloc((void**)&buf.p, sizeof(double) * chunk,
                         hipHostMallocDefault));
  for (int k = 0; k < *nfields; ++k) {
    double* dst = (double*)fields[k];
    for (int64_t o = 0; o < *n; o += chunk) {
      const int64_t m = *n - o < chunk ? *n - o : chunk;
      if (std::fread(buf.p, sizeof(double), (size_t)m, file.f) != (size_t)m)
        return fail_arg(__func__, "short checkpoint file");
      NK_HIP(hipMemcpyAsync(dst + o, buf.p, sizeof(double) * m,
                            hipMemcpyHostToDevice, stream()));
      NK_HIP(hipStreamSynchronize(stream()));
    }
  }
  *tstep = h.tstep;
  *time = h.time;
  *dt = h.dt;
  return kOk;
}

}  // extern "C"
