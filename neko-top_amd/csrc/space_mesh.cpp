// space_mesh.cpp -- host-side set-up of the SEM discretisation: GLL space,
// the synthetic box mesh with seeded vertex jitter, dofmap global ids,
// Dirichlet face masks and the canonical gather-scatter map.
//
// These produce the inputs of the operator core (what Neko's space_t,
// dofmap_t and gs_t%init build before the reference calls ax_t / gs_t /
// ksp_t, e.g. sources/adjoint/adjoint_scheme.f90:331-343).  Conventions:
//   local point  l = i + lx*(j + lx*(k + lx*e)), 0-based, i fastest
//   element      e = ex + nx*(ey + ny*ez)
//   global id    g = gx + (nx*p+1)*(gy + (ny*p+1)*gz), gx = ex*p + i
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"
#include "nekoamd.h"

using namespace nekoamd;

namespace {

// Counter-based generator: SplitMix64 finaliser applied to seed + (idx+1)*phi.
inline uint64_t mix64(uint64_t seed, uint64_t idx) {
  uint64_t z = seed + (idx + 1u) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline double unit_uniform(uint64_t seed, uint64_t idx) {
  return (double)(mix64(seed, idx) >> 11) * 0x1.0p-53;
}

// P_N(x) and P_{N-1}(x) by the three-term recurrence.
void legendre_pair(int N, double x, double& pn, double& pnm1) {
  double a = 1.0, b = x;
  if (N == 0) {
    pn = 1.0;
    pnm1 = 0.0;
    return;
  }
  for (int k = 2; k <= N; ++k) {
    const double c = ((2.0 * k - 1.0) * x * b - (k - 1.0) * a) / k;
    a = b;
    b = c;
  }
  pn = b;
  pnm1 = a;
}

}  // namespace

extern "C" {

uint64_t nekoamd_rng_u64(const uint64_t* seed, const uint64_t* idx) {
  return mix64(*seed, *idx);
}

int nekoamd_gll(const int* lx, double* z, double* w, double* dx) {
  NK_REQUIRE(lx && z && w && dx, "null argument");
  const int n = *lx, N = n - 1;
  NK_REQUIRE(n >= 2 && n <= NEKOAMD_LX_MAX, "lx out of range");
  // Nodes: Newton iteration on x P_N - P_{N-1} (zero at the interior GLL
  // nodes) from the Chebyshev-Lobatto points, made exactly antisymmetric.
  for (int i = 0; 2 * i <= N; ++i) {
    double x;
    if (i == 0) {
      x = -1.0;
    } else if (2 * i == N) {
      x = 0.0;
    } else {
      x = -std::cos(M_PI * i / N);
      for (int it = 0; it < 100; ++it) {
        double pn, pm;
        legendre_pair(N, x, pn, pm);
        const double step = (x * pn - pm) / ((N + 1.0) * pn);
        x -= step;
        if (std::fabs(step) < 1e-17) break;
      }
    }
    z[i] = x;
    z[N - i] = -x;
  }
  std::vector<double> pn(n);
  for (int i = 0; i < n; ++i) {
    double pm;
    legendre_pair(N, z[i], pn[i], pm);
    w[i] = 2.0 / (N * (N + 1.0) * pn[i] * pn[i]);
  }
  // dx[i*lx + j] = l_j'(z_i)
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double v = 0.0;
      if (i != j)
        v = pn[i] / (pn[j] * (z[i] - z[j]));
      else if (i == 0)
        v = -N * (N + 1.0) / 4.0;
      else if (i == N)
        v = N * (N + 1.0) / 4.0;
      dx[i * n + j] = v;
    }
  return kOk;
}

int nekoamd_box_vertices(const int* nx, const int* ny, const int* nz,
                         const double* lxyz, const double* jitter,
                         const uint64_t* seed, double* vxyz) {
  NK_REQUIRE(nx && ny && nz && lxyz && jitter && seed && vxyz, "null argument");
  NK_REQUIRE(*nx > 0 && *ny > 0 && *nz > 0, "empty mesh");
  const int n[3] = {*nx, *ny, *nz};
  const double h[3] = {lxyz[0] / n[0], lxyz[1] / n[1], lxyz[2] / n[2]};
  const int64_t sx = n[0] + 1, sy = n[1] + 1;
  for (int vz = 0; vz <= n[2]; ++vz)
    for (int vy = 0; vy <= n[1]; ++vy)
      for (int vx = 0; vx <= n[0]; ++vx) {
        const int64_t v = vx + sx * (vy + sy * vz);
        const int q[3] = {vx, vy, vz};
        for (int a = 0; a < 3; ++a) {
          const double r = unit_uniform(*seed, 3u * (uint64_t)v + a);
          double c = q[a] * h[a];
          if (q[a] > 0 && q[a] < n[a]) c += *jitter * h[a] * (2.0 * r - 1.0);
          vxyz[3 * v + a] = c;
        }
      }
  return kOk;
}

int nekoamd_box_corners(const int* nx, const int* ny, const int* nz,
                        const double* vxyz, double* corners) {
  NK_REQUIRE(nx && ny && nz && vxyz && corners, "null argument");
  const int64_t sx = *nx + 1, sy = *ny + 1;
  int64_t e = 0;
  for (int ez = 0; ez < *nz; ++ez)
    for (int ey = 0; ey < *ny; ++ey)
      for (int ex = 0; ex < *nx; ++ex, ++e)
        for (int c = 0; c < 8; ++c) {
          const int64_t v = (ex + (c & 1)) + sx * ((ey + ((c >> 1) & 1)) +
                                                   sy * (ez + ((c >> 2) & 1)));
          std::memcpy(corners + (e * 8 + c) * 3, vxyz + 3 * v,
                      3 * sizeof(double));
        }
  return kOk;
}

int nekoamd_box_gids(const int* nx, const int* ny, const int* nz,
                     const int* lx, int64_t* gid) {
  NK_REQUIRE(nx && ny && nz && lx && gid, "null argument");
  const int p = *lx - 1, L = *lx;
  const int64_t sx = (int64_t)*nx * p + 1, sy = (int64_t)*ny * p + 1;
  int64_t l = 0;
  for (int ez = 0; ez < *nz; ++ez)
    for (int ey = 0; ey < *ny; ++ey)
      for (int ex = 0; ex < *nx; ++ex)
        for (int k = 0; k < L; ++k)
          for (int j = 0; j < L; ++j) {
            const int64_t row =
                sx * (((int64_t)ey * p + j) + sy * ((int64_t)ez * p + k));
            for (int i = 0; i < L; ++i) gid[l++] = (int64_t)ex * p + i + row;
          }
  return kOk;
}

int nekoamd_box_face_mask(const int* nx, const int* ny, const int* nz,
                          const int* lx, const int* faces, int32_t* out,
                          int64_t* count) {
  NK_REQUIRE(nx && ny && nz && lx && faces && count, "null argument");
  const int L = *lx, p = L - 1, f = *faces;
  int64_t cnt = 0, l = 0;
  for (int ez = 0; ez < *nz; ++ez)
    for (int ey = 0; ey < *ny; ++ey)
      for (int ex = 0; ex < *nx; ++ex)
        for (int k = 0; k < L; ++k)
          for (int j = 0; j < L; ++j)
            for (int i = 0; i < L; ++i, ++l) {
              const bool on = ((f & NEKOAMD_FACE_XMIN) && ex == 0 && i == 0) ||
                              ((f & NEKOAMD_FACE_XMAX) && ex == *nx - 1 && i == p) ||
                              ((f & NEKOAMD_FACE_YMIN) && ey == 0 && j == 0) ||
                              ((f & NEKOAMD_FACE_YMAX) && ey == *ny - 1 && j == p) ||
                              ((f & NEKOAMD_FACE_ZMIN) && ez == 0 && k == 0) ||
                              ((f & NEKOAMD_FACE_ZMAX) && ez == *nz - 1 && k == p);
              if (on) {
                if (out) out[cnt] = (int32_t)l;
                ++cnt;
              }
            }
  *count = cnt;
  return kOk;
}

}  // extern "C"

namespace nekoamd {

// Canonical bucketed gather-scatter map.  Buckets by multiplicity m
// ascending; within a bucket tuples by global id ascending; within a tuple
// local indices ascending.  Points of multiplicity one are left out (gs
// leaves them untouched).
int build_gs_map(int64_t n, const int64_t* gid, std::vector<int>& bm,
                 std::vector<int64_t>& bcnt, std::vector<int32_t>& idx) {
  bm.clear();
  bcnt.clear();
  idx.clear();
  if (n == 0) return kOk;
  if (n > INT32_MAX) return fail_arg("gs_init", "more than 2^31-1 local points");
  int64_t gmax = 0;
  for (int64_t l = 0; l < n; ++l) {
    if (gid[l] < 0) return fail_arg("gs_init", "negative global id");
    gmax = std::max(gmax, gid[l]);
  }
  // (gid, l) pairs in ascending order: a counting sort when the ids are
  // dense, else a comparison sort.
  std::vector<int32_t> order(n);
  std::vector<int64_t> gsorted;
  if (gmax <= 4 * n + 1024) {
    std::vector<int64_t> start(gmax + 2, 0);
    for (int64_t l = 0; l < n; ++l) start[gid[l] + 1]++;
    for (int64_t g = 0; g <= gmax; ++g) start[g + 1] += start[g];
    std::vector<int64_t> pos(start.begin(), start.end() - 1);
    for (int64_t l = 0; l < n; ++l) order[pos[gid[l]]++] = (int32_t)l;
  } else {
    for (int64_t l = 0; l < n; ++l) order[l] = (int32_t)l;
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return gid[a] < gid[b]; });
  }
  // Segment lengths in id order.
  std::vector<int64_t> seg_begin;
  std::vector<int> seg_len;
  for (int64_t s = 0; s < n;) {
    int64_t t = s + 1;
    while (t < n && gid[order[t]] == gid[order[s]]) ++t;
    if (t - s >= 2) {
      if (t - s > NEKOAMD_GS_MAX_MULT)
        return fail_arg("gs_init", "multiplicity above NEKOAMD_GS_MAX_MULT");
      seg_begin.push_back(s);
      seg_len.push_back((int)(t - s));
    }
    s = t;
  }
  std::vector<int64_t> cnt(NEKOAMD_GS_MAX_MULT + 1, 0);
  for (int m : seg_len) cnt[m]++;
  std::vector<int64_t> off(NEKOAMD_GS_MAX_MULT + 1, 0);
  int64_t tot = 0;
  for (int m = 2; m <= NEKOAMD_GS_MAX_MULT; ++m) {
    off[m] = tot;
    tot += cnt[m] * m;
    if (cnt[m]) {
      bm.push_back(m);
      bcnt.push_back(cnt[m]);
    }
  }
  idx.resize(tot);
  for (size_t s = 0; s < seg_len.size(); ++s) {
    const int m = seg_len[s];
    for (int q = 0; q < m; ++q) idx[off[m] + q] = order[seg_begin[s] + q];
    off[m] += m;
  }
  return kOk;
}

}  // namespace nekoamd
