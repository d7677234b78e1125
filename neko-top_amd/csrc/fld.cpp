// fld.cpp -- field output in Neko's fld format (the Nek5000 binary field
// file), as the reference samples it through Neko's fld_file_output_t:
// the optimisation output of steady_state_problem.f90:228-244 (13 fields:
// p, u, v, w, design -> "temperature", adjoint u, v, w, p, chi and the
// sensitivities -> scalars s1..s8, :216-226) and the design output of
// topopt_design.f90:283-286.
//
// One file, little endian, one rank:
//   132 ASCII bytes   #std wdsize lx ly lz nelv nelgt time step fid0 nfiles
//                     rdcode, blank padded (Fortran format '#std',1x,i1,1x,
//                     i2,1x,i2,1x,i2,1x,i10,1x,i10,1x,e20.13,1x,i9,1x,i6,1x,
//                     i6,1x,<rdcode>)
//   float32 6.54321   byte-order tag
//   int32 [nelv]      element numbers, 1-based
//   groups in rdcode order, element blocks of lx^3 words of wdsize bytes:
//     X  x, y, z of each element      U  u, v, w of each element
//     P  p of each element            T  t of each element
//     S  nn scalars: all elements of scalar 1, then scalar 2, ...
//   float32 (min, max) per element and component, group by group (3-D
//   metadata: X and U as xmin xmax ymin ymax zmin zmax per element)
// A list of n fields is written as Neko's fld_file does: 1 -> P, 2 -> P T,
// 3 -> U, 4 -> P U, 5 -> P U T, n > 5 -> P U T S(n - 5), taking the fields
// in list order p, u, v, w, t, s1, ...; the mesh (X) goes into the first
// file of a series only.  The series index name.nek5000 has the three
// lines of examples/power_iterations/baseflow0.nek5000.
//
// Host code: device fields are copied to the host first (device = 1).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "internal.h"
#include "nekoamd.h"

namespace nekoamd {
namespace {

// Fortran Ew.d with w = 20, d = 13: [-]0.ddddddddddddd E+xx, right aligned.
std::string fortran_e20_13(double v) {
  char mant[32];
  int ex = 0;
  const bool neg = v < 0.0;
  double a = std::fabs(v);
  long long digits = 0;
  if (a != 0.0) {
    ex = (int)std::floor(std::log10(a)) + 1;
    double m = a / std::pow(10.0, ex);
    if (m >= 1.0) {
      m /= 10.0;
      ++ex;
    } else if (m < 0.1) {
      m *= 10.0;
      --ex;
    }
    digits = std::llround(m * 1e13);
    if (digits >= 10000000000000LL) {  // rounded up to 1.0
      digits /= 10;
      ++ex;
    }
  }
  std::snprintf(mant, sizeof(mant), "%s0.%013lldE%c%02d", neg ? "-" : "",
                digits, ex < 0 ? '-' : '+', std::abs(ex));
  std::string s(mant);
  if (s.size() < 20) s.insert(0, 20 - s.size(), ' ');
  return s;
}

template <typename W>
void put(std::FILE* f, const W* a, size_t n, size_t& bad) {
  if (std::fwrite(a, sizeof(W), n, f) != n) ++bad;
}

}  // namespace
}  // namespace nekoamd

using namespace nekoamd;

extern "C" {

int nekoamd_fld_write(const char* path, const double* time, const int* step,
                      const int* lx, const int* nelv, const int* nfields,
                      const void* const* fields, const void* x, const void* y,
                      const void* z, const int* wdsize, const int* device) {
  NK_REQUIRE(path && time && step && lx && nelv && nfields && wdsize && device,
             "null argument");
  const int nf = *nfields, L = *lx, ne = *nelv, ws = *wdsize;
  NK_REQUIRE(L >= 2 && L <= 99, "lx out of range");
  NK_REQUIRE(ne >= 1, "nelv < 1");
  NK_REQUIRE(nf >= 0 && (nf == 0 || fields != nullptr), "bad field list");
  NK_REQUIRE(ws == 4 || ws == 8, "wdsize must be 4 or 8");
  const bool mesh = x != nullptr;
  NK_REQUIRE(!mesh || (y && z), "x given without y, z");
  NK_REQUIRE(mesh || nf > 0, "nothing to write");
  for (int i = 0; i < nf; ++i) NK_REQUIRE(fields[i] != nullptr, "null field");
  if (*device) NK_INIT();
  const size_t np = (size_t)L * L * L, n = np * ne;

  // fetch everything as host doubles, in list order
  auto fetch = [&](const void* p, std::vector<double>& h) -> int {
    h.resize(n);
    if (*device) {
      hipError_t e = hipMemcpy(h.data(), p, n * sizeof(double),
                               hipMemcpyDeviceToHost);
      if (e != hipSuccess) return fail_hip("nekoamd_fld_write", e);
    } else {
      std::memcpy(h.data(), p, n * sizeof(double));
    }
    return kOk;
  };
  if (*device) NK_HIP(hipStreamSynchronize(stream()));
  std::vector<std::vector<double>> xyz(mesh ? 3 : 0), fl(nf);
  for (int d = 0; d < (mesh ? 3 : 0); ++d) {
    const int rc = fetch(d == 0 ? x : d == 1 ? y : z, xyz[d]);
    if (rc != kOk) return rc;
  }
  for (int i = 0; i < nf; ++i) {
    const int rc = fetch(fields[i], fl[i]);
    if (rc != kOk) return rc;
  }

  // groups: (code, components as indices into fl)
  struct Group {
    char code;
    std::vector<const std::vector<double>*> comp;
  };
  std::vector<Group> groups;
  if (mesh) groups.push_back({'X', {&xyz[0], &xyz[1], &xyz[2]}});
  int ns = 0;
  switch (nf) {
    case 0: break;
    case 1: groups.push_back({'P', {&fl[0]}}); break;
    case 2:
      groups.push_back({'P', {&fl[0]}});
      groups.push_back({'T', {&fl[1]}});
      break;
    case 3: groups.push_back({'U', {&fl[0], &fl[1], &fl[2]}}); break;
    default:
      groups.push_back({'P', {&fl[0]}});
      groups.push_back({'U', {&fl[1], &fl[2], &fl[3]}});
      if (nf >= 5) groups.push_back({'T', {&fl[4]}});
      if (nf > 5) {
        Group s{'S', {}};
        for (int i = 5; i < nf; ++i) s.comp.push_back(&fl[i]);
        ns = nf - 5;
        groups.push_back(s);
      }
  }
  std::string rdcode;
  for (const Group& g : groups) rdcode += g.code;
  if (ns > 0) {
    char b[8];
    std::snprintf(b, sizeof(b), "%02d", ns);
    rdcode += b;
  }
  NK_REQUIRE(ns <= 99, "too many scalars");

  char head[160];
  std::snprintf(head, sizeof(head), "#std %1d %2d %2d %2d %10d %10d %s %9d %6d %6d %s",
                ws, L, L, L, ne, ne, fortran_e20_13(*time).c_str(), *step, 0,
                1, rdcode.c_str());
  std::string hdr(head);
  NK_REQUIRE(hdr.size() <= 132, "header too long");
  hdr.resize(132, ' ');

  std::FILE* f = std::fopen(path, "wb");
  if (!f) {
    set_error("nekoamd_fld_write", (std::string("cannot open ") + path).c_str());
    return kErrArg;
  }
  size_t bad = 0;
  put(f, hdr.data(), hdr.size(), bad);
  const float tag = 6.54321f;
  put(f, &tag, 1, bad);
  std::vector<int32_t> ids(ne);
  for (int e = 0; e < ne; ++e) ids[e] = e + 1;
  put(f, ids.data(), ids.size(), bad);

  std::vector<float> b4(np);
  auto block = [&](const std::vector<double>& a, int e) {
    const double* src = a.data() + (size_t)e * np;
    if (ws == 8) {
      put(f, src, np, bad);
    } else {
      for (size_t i = 0; i < np; ++i) b4[i] = (float)src[i];
      put(f, b4.data(), np, bad);
    }
  };
  for (const Group& g : groups) {
    if (g.code == 'S') {
      for (const auto* c : g.comp)
        for (int e = 0; e < ne; ++e) block(*c, e);
    } else {
      for (int e = 0; e < ne; ++e)
        for (const auto* c : g.comp) block(*c, e);
    }
  }
  // metadata: per element (min, max) of every component, as float32
  std::vector<float> md;
  for (const Group& g : groups) {
    auto mm = [&](const std::vector<double>& a, int e) {
      const double* s = a.data() + (size_t)e * np;
      const auto r = std::minmax_element(s, s + np);
      md.push_back((float)*r.first);
      md.push_back((float)*r.second);
    };
    if (g.code == 'S') {
      for (const auto* c : g.comp)
        for (int e = 0; e < ne; ++e) mm(*c, e);
    } else {
      for (int e = 0; e < ne; ++e)
        for (const auto* c : g.comp) mm(*c, e);
    }
  }
  put(f, md.data(), md.size(), bad);
  if (std::fclose(f) != 0) ++bad;
  if (bad) {
    set_error("nekoamd_fld_write", "short write");
    return kErrArg;
  }
  return kOk;
}

int nekoamd_fld_series(const char* path, const char* prefix,
                       const int* first, const int* count) {
  NK_REQUIRE(path && prefix && first && count, "null argument");
  std::FILE* f = std::fopen(path, "w");
  if (!f) {
    set_error("nekoamd_fld_series", (std::string("cannot open ") + path).c_str());
    return kErrArg;
  }
  const int w = std::fprintf(f, "filetemplate: %s%%01d.f%%05d\nfirsttimestep: %d\nnumtimesteps: %d\n",
                             prefix, *first, *count);
  if (std::fclose(f) != 0 || w < 0) {
    set_error("nekoamd_fld_series", "short write");
    return kErrArg;
  }
  return kOk;
}

}  // extern "C"
