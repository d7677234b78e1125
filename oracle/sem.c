/*
 * sem.c -- FP64 CPU restatement of the spectral-element operator core:
 * GLL space, synthetic box mesh, dofmap ids, geometric factors, ax_helm,
 * gather-scatter, Jacobi and preconditioned CG.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle).  The product library never links or
 * calls this file.
 *
 * These algorithms live in Neko, which neko-top builds on but does not ship
 * (CMakeLists.txt:74, :144-157 of the reference require an external
 * NEKO_DIR).  They are restated here from the textbook SEM formulation
 * (Deville, Fischer & Mund 2002, ch. 4) using Neko's conventions:
 *   - ax_helm  w = sum_a D_a^T h1 sum_b G_ab D_b u  [+ h2 B u]   (factory
 *     call sites: sources/adjoint/adjoint_pnpn.f90:207, :217;
 *     sources/mapping_functions/PDE_filter_mapping.f90:177)
 *   - gs_op ADD = Q Q^T  (sources/adjoint/adjoint_scheme.f90:341,
 *     sources/adjoint/adjoint_pnpn.f90:725, :755-757)
 *   - Jacobi diag without the G_ab cross terms, then gs ADD and invert
 *     (sources/adjoint/adjoint_scheme.f90:1058-1092)
 *   - CG with mult-weighted inner products and ||r|| = sqrt(<r,r>)/sqrt(vol)
 *     (sources/adjoint/adjoint_pnpn.f90:734-736, adjoint_scheme.f90:1046-1055)
 * None of these is pinned by the reference itself; tests/ pin them with
 * closed forms and dense brute-force assembly.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
static int g_threads = 1;

void or_set_threads(int nthreads) {
  g_threads = nthreads < 1 ? 1 : nthreads;
}
int or_get_threads(void) { return g_threads; }

/* ------------------------------------------------------------------------ */
uint64_t or_splitmix64(uint64_t seed, uint64_t idx) {
  uint64_t z = seed + (idx + 1u) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double or_u01(uint64_t seed, uint64_t idx) {
  return (double)(or_splitmix64(seed, idx) >> 11) * 0x1.0p-53;
}

/* ------------------------------------------------------------------------ */
static void legendre2(int N, double x, double *PN, double *PNm1) {
  double p0 = 1.0, p1 = x;
  if (N == 0) {
    *PN = 1.0;
    *PNm1 = 0.0;
    return;
  }
  for (int k = 2; k <= N; ++k) {
    double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  *PN = p1;
  *PNm1 = p0;
}

/* Gauss-Lobatto-Legendre nodes z (ascending), weights w and derivative
 * matrix D[i*lx+j] = l_j'(z_i).  Nodes: Newton on x P_N - P_{N-1} = 0
 * started from the Chebyshev-Lobatto points; mirrored to be exactly
 * symmetric. */
void or_gll(int lx, double *z, double *w, double *D) {
  const int N = lx - 1;
  for (int i = 0; i <= N / 2; ++i) {
    double x = -cos(M_PI * i / N);
    if (i == 0) {
      x = -1.0;
    } else if (2 * i == N) {
      x = 0.0;
    } else {
      for (int it = 0; it < 100; ++it) {
        double PN, PNm1;
        legendre2(N, x, &PN, &PNm1);
        double dx = (x * PN - PNm1) / ((N + 1.0) * PN);
        x -= dx;
        if (fabs(dx) < 1e-17) break;
      }
    }
    z[i] = x;
    z[N - i] = -x;
  }
  for (int i = 0; i <= N; ++i) {
    double PN, PNm1;
    legendre2(N, z[i], &PN, &PNm1);
    w[i] = 2.0 / (N * (N + 1.0) * PN * PN);
  }
  for (int i = 0; i <= N; ++i) {
    double Pi, t;
    legendre2(N, z[i], &Pi, &t);
    for (int j = 0; j <= N; ++j) {
      double Pj;
      legendre2(N, z[j], &Pj, &t);
      double v;
      if (i != j)
        v = Pi / (Pj * (z[i] - z[j]));
      else if (i == 0)
        v = -N * (N + 1.0) / 4.0;
      else if (i == N)
        v = N * (N + 1.0) / 4.0;
      else
        v = 0.0;
      D[i * lx + j] = v;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Box [0,Lx]x[0,Ly]x[0,Lz] of nx*ny*nz hexes.  A vertex coordinate is
 * jittered by U(-jitter,jitter)*h along an axis iff the vertex is interior
 * along that axis (boundary vertices move only tangentially).  RNG stream:
 * or_u01(seed, 3*v + axis). */
void or_box_vertices(int nx, int ny, int nz, double Lx, double Ly, double Lz,
                     double jitter, uint64_t seed, double *vxyz) {
  const double hx = Lx / nx, hy = Ly / ny, hz = Lz / nz;
  const int64_t nvx = nx + 1, nvy = ny + 1;
  for (int vz = 0; vz <= nz; ++vz)
    for (int vy = 0; vy <= ny; ++vy)
      for (int vx = 0; vx <= nx; ++vx) {
        int64_t v = vx + nvx * (vy + nvy * vz);
        double c[3] = {vx * hx, vy * hy, vz * hz};
        const int inner[3] = {vx > 0 && vx < nx, vy > 0 && vy < ny,
                              vz > 0 && vz < nz};
        const double h[3] = {hx, hy, hz};
        for (int a = 0; a < 3; ++a) {
          double r = or_u01(seed, 3u * (uint64_t)v + a);
          if (inner[a]) c[a] += jitter * h[a] * (2.0 * r - 1.0);
          vxyz[3 * v + a] = c[a];
        }
      }
}

void or_element_corners(int nx, int ny, int nz, const double *vxyz,
                        double *corners) {
  const int64_t nvx = nx + 1, nvy = ny + 1;
  for (int ez = 0; ez < nz; ++ez)
    for (int ey = 0; ey < ny; ++ey)
      for (int ex = 0; ex < nx; ++ex) {
        int64_t e = ex + (int64_t)nx * (ey + (int64_t)ny * ez);
        for (int c = 0; c < 8; ++c) {
          int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
          int64_t v = (ex + cx) + nvx * ((ey + cy) + nvy * (ez + cz));
          for (int a = 0; a < 3; ++a)
            corners[(e * 8 + c) * 3 + a] = vxyz[3 * v + a];
        }
      }
}

void or_dofmap_gids(int nx, int ny, int nz, int lx, int64_t *gid) {
  const int p = lx - 1;
  const int64_t ngx = (int64_t)nx * p + 1, ngy = (int64_t)ny * p + 1;
  int64_t l = 0;
  for (int ez = 0; ez < nz; ++ez)
    for (int ey = 0; ey < ny; ++ey)
      for (int ex = 0; ex < nx; ++ex)
        for (int k = 0; k < lx; ++k)
          for (int j = 0; j < lx; ++j)
            for (int i = 0; i < lx; ++i) {
              int64_t gx = (int64_t)ex * p + i, gy = (int64_t)ey * p + j,
                      gz = (int64_t)ez * p + k;
              gid[l++] = gx + ngx * (gy + ngy * gz);
            }
  (void)ny;
}

int64_t or_face_mask(int nx, int ny, int nz, int lx, int faces,
                     int32_t *out) {
  const int p = lx - 1;
  int64_t cnt = 0, l = 0;
  for (int ez = 0; ez < nz; ++ez)
    for (int ey = 0; ey < ny; ++ey)
      for (int ex = 0; ex < nx; ++ex)
        for (int k = 0; k < lx; ++k)
          for (int j = 0; j < lx; ++j)
            for (int i = 0; i < lx; ++i, ++l) {
              int on = ((faces & 1) && ex == 0 && i == 0) ||
                       ((faces & 2) && ex == nx - 1 && i == p) ||
                       ((faces & 4) && ey == 0 && j == 0) ||
                       ((faces & 8) && ey == ny - 1 && j == p) ||
                       ((faces & 16) && ez == 0 && k == 0) ||
                       ((faces & 32) && ez == nz - 1 && k == p);
              if (on) {
                if (out) out[cnt] = (int32_t)l;
                ++cnt;
              }
            }
  return cnt;
}

/* ------------------------------------------------------------------------ */
/* Geometric factors of Neko's coef_t: GLL coordinates by the trilinear map
 * of the 8 corners, x_r = D x (spectral derivative), cofactors
 * c_r = x_s x x_t, c_s = x_t x x_r, c_t = x_r x x_s, J = x_r . c_r,
 * G_ab = w_i w_j w_k (c_a . c_b) / J,  B = w_i w_j w_k J. */
void or_coef(int lx, int nelv, const double *corners, const double *z,
             const double *w, const double *D, double *x, double *y,
             double *zz, double *G11, double *G12, double *G13, double *G22,
             double *G23, double *G33, double *B, double *jac) {
  const int np = lx * lx * lx;
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (int e = 0; e < nelv; ++e) {
    const double *X = corners + (int64_t)e * 24;
    double *xe = x + (int64_t)e * np, *ye = y + (int64_t)e * np,
           *ze = zz + (int64_t)e * np;
    for (int k = 0; k < lx; ++k)
      for (int j = 0; j < lx; ++j)
        for (int i = 0; i < lx; ++i) {
          double a0 = 0.5 * (1.0 - z[i]), a1 = 0.5 * (1.0 + z[i]);
          double b0 = 0.5 * (1.0 - z[j]), b1 = 0.5 * (1.0 + z[j]);
          double c0 = 0.5 * (1.0 - z[k]), c1 = 0.5 * (1.0 + z[k]);
          int l = i + lx * (j + lx * k);
          double out[3];
          for (int a = 0; a < 3; ++a) {
            double e00 = a0 * X[0 * 3 + a] + a1 * X[1 * 3 + a];
            double e10 = a0 * X[2 * 3 + a] + a1 * X[3 * 3 + a];
            double e01 = a0 * X[4 * 3 + a] + a1 * X[5 * 3 + a];
            double e11 = a0 * X[6 * 3 + a] + a1 * X[7 * 3 + a];
            out[a] = c0 * (b0 * e00 + b1 * e10) + c1 * (b0 * e01 + b1 * e11);
          }
          xe[l] = out[0];
          ye[l] = out[1];
          ze[l] = out[2];
        }
    for (int k = 0; k < lx; ++k)
      for (int j = 0; j < lx; ++j)
        for (int i = 0; i < lx; ++i) {
          int l = i + lx * (j + lx * k);
          double xr[3] = {0, 0, 0}, xs[3] = {0, 0, 0}, xt[3] = {0, 0, 0};
          const double *f[3] = {xe, ye, ze};
          for (int a = 0; a < 3; ++a)
            for (int m = 0; m < lx; ++m) {
              xr[a] += D[i * lx + m] * f[a][m + lx * (j + lx * k)];
              xs[a] += D[j * lx + m] * f[a][i + lx * (m + lx * k)];
              xt[a] += D[k * lx + m] * f[a][i + lx * (j + lx * m)];
            }
          double cr[3] = {xs[1] * xt[2] - xs[2] * xt[1],
                          xs[2] * xt[0] - xs[0] * xt[2],
                          xs[0] * xt[1] - xs[1] * xt[0]};
          double cs[3] = {xt[1] * xr[2] - xt[2] * xr[1],
                          xt[2] * xr[0] - xt[0] * xr[2],
                          xt[0] * xr[1] - xt[1] * xr[0]};
          double ct[3] = {xr[1] * xs[2] - xr[2] * xs[1],
                          xr[2] * xs[0] - xr[0] * xs[2],
                          xr[0] * xs[1] - xr[1] * xs[0]};
          double J = xr[0] * cr[0] + xr[1] * cr[1] + xr[2] * cr[2];
          double w3 = w[i] * w[j] * w[k];
          double s = w3 / J;
          int64_t gl = (int64_t)e * np + l;
          G11[gl] = s * (cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
          G12[gl] = s * (cr[0] * cs[0] + cr[1] * cs[1] + cr[2] * cs[2]);
          G13[gl] = s * (cr[0] * ct[0] + cr[1] * ct[1] + cr[2] * ct[2]);
          G22[gl] = s * (cs[0] * cs[0] + cs[1] * cs[1] + cs[2] * cs[2]);
          G23[gl] = s * (cs[0] * ct[0] + cs[1] * ct[1] + cs[2] * ct[2]);
          G33[gl] = s * (ct[0] * ct[0] + ct[1] * ct[1] + ct[2] * ct[2]);
          B[gl] = w3 * J;
          if (jac) jac[gl] = J;
        }
  }
}

/* ------------------------------------------------------------------------ */
void or_ax_helm(int lx, int nelv, double *w, const double *u, const double *D,
                const double *h1, const double *h2, const double *G11,
                const double *G12, const double *G13, const double *G22,
                const double *G23, const double *G33, const double *B,
                int ifh2) {
  const int np = lx * lx * lx;
#pragma omp parallel num_threads(g_threads)
  {
    double *wr = (double *)malloc(sizeof(double) * 3 * np);
    double *ws = wr + np, *wt = wr + 2 * np;
#pragma omp for schedule(static)
    for (int e = 0; e < nelv; ++e) {
      const int64_t o = (int64_t)e * np;
      const double *ue = u + o;
      for (int k = 0; k < lx; ++k)
        for (int j = 0; j < lx; ++j)
          for (int i = 0; i < lx; ++i) {
            double ur = 0, us = 0, ut = 0;
            for (int l = 0; l < lx; ++l) {
              ur += D[i * lx + l] * ue[l + lx * (j + lx * k)];
              us += D[j * lx + l] * ue[i + lx * (l + lx * k)];
              ut += D[k * lx + l] * ue[i + lx * (j + lx * l)];
            }
            int p = i + lx * (j + lx * k);
            int64_t g = o + p;
            double h = h1[g];
            wr[p] = h * (G11[g] * ur + G12[g] * us + G13[g] * ut);
            ws[p] = h * (G12[g] * ur + G22[g] * us + G23[g] * ut);
            wt[p] = h * (G13[g] * ur + G23[g] * us + G33[g] * ut);
          }
      for (int k = 0; k < lx; ++k)
        for (int j = 0; j < lx; ++j)
          for (int i = 0; i < lx; ++i) {
            double s = 0;
            for (int l = 0; l < lx; ++l) {
              s += D[l * lx + i] * wr[l + lx * (j + lx * k)];
              s += D[l * lx + j] * ws[i + lx * (l + lx * k)];
              s += D[l * lx + k] * wt[i + lx * (j + lx * l)];
            }
            int64_t g = o + i + lx * (j + lx * k);
            if (ifh2) s += h2[g] * B[g] * ue[i + lx * (j + lx * k)];
            w[g] = s;
          }
    }
    free(wr);
  }
}

/* ------------------------------------------------------------------------ */
/* Canonical bucketed gather-scatter map (counting sort on the global id). */
int64_t or_gs_build(int64_t n, const int64_t *gid, int32_t *idx, int *nb,
                    int *bm, int64_t *bcnt) {
  int64_t ng = 0;
  for (int64_t l = 0; l < n; ++l)
    if (gid[l] + 1 > ng) ng = gid[l] + 1;
  int64_t *start = (int64_t *)calloc((size_t)ng + 1, sizeof(int64_t));
  for (int64_t l = 0; l < n; ++l) start[gid[l] + 1]++;
  int64_t cnt_m[OR_GS_MAXM + 1];
  memset(cnt_m, 0, sizeof(cnt_m));
  for (int64_t g = 0; g < ng; ++g) {
    int64_t m = start[g + 1];
    if (m >= 2 && m <= OR_GS_MAXM) cnt_m[m]++;
  }
  for (int64_t g = 0; g < ng; ++g) start[g + 1] += start[g];
  int32_t *order = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
  int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)ng + 8);
  memcpy(fill, start, sizeof(int64_t) * (size_t)ng);
  for (int64_t l = 0; l < n; ++l) order[fill[gid[l]]++] = (int32_t)l;
  int64_t off_m[OR_GS_MAXM + 1];
  int64_t tot = 0;
  int b = 0;
  for (int m = 2; m <= OR_GS_MAXM; ++m) {
    off_m[m] = tot;
    tot += cnt_m[m] * m;
    if (cnt_m[m]) {
      if (bm) bm[b] = m;
      if (bcnt) bcnt[b] = cnt_m[m];
      ++b;
    }
  }
  if (nb) *nb = b;
  if (idx) {
    for (int64_t g = 0; g < ng; ++g) {
      int64_t m = start[g + 1] - start[g];
      if (m < 2 || m > OR_GS_MAXM) continue;
      for (int64_t q = 0; q < m; ++q) idx[off_m[m] + q] = order[start[g] + q];
      off_m[m] += m;
    }
  }
  free(order);
  free(fill);
  free(start);
  return tot;
}

void or_gs_op(double *u, int nb, const int *bm, const int64_t *bcnt,
              const int32_t *idx, int op) {
  int64_t off = 0;
  for (int b = 0; b < nb; ++b) {
    const int m = bm[b];
    for (int64_t t = 0; t < bcnt[b]; ++t) {
      const int32_t *q = idx + off + t * m;
      double s = u[q[0]];
      for (int r = 1; r < m; ++r) {
        double v = u[q[r]];
        switch (op) {
          case OR_GS_ADD: s = s + v; break;
          case OR_GS_MUL: s = s * v; break;
          case OR_GS_MIN: s = v < s ? v : s; break;
          default: s = v > s ? v : s; break;
        }
      }
      for (int r = 0; r < m; ++r) u[q[r]] = s;
    }
    off += bcnt[b] * m;
  }
}

/* ------------------------------------------------------------------------ */
/* Deterministic reductions: fixed static chunks per thread, partials summed
 * in thread order. */
double or_glsc3(int64_t n, const double *a, const double *b, const double *c) {
  int T = g_threads;
  double part[256];
  if (T > 256) T = 256;
#pragma omp parallel num_threads(T)
  {
    int t = 0, nt = 1;
#ifdef _OPENMP
    t = omp_get_thread_num();
    nt = omp_get_num_threads();
#endif
    int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
    double s = 0.0;
    for (int64_t i = lo; i < hi; ++i) s = s + a[i] * b[i] * c[i];
    part[t] = s;
    if (t == 0) T = nt;
  }
  double s = 0.0;
  for (int t = 0; t < T; ++t) s += part[t];
  return s;
}

double or_glsc2(int64_t n, const double *a, const double *b) {
  int T = g_threads;
  double part[256];
  if (T > 256) T = 256;
#pragma omp parallel num_threads(T)
  {
    int t = 0, nt = 1;
#ifdef _OPENMP
    t = omp_get_thread_num();
    nt = omp_get_num_threads();
#endif
    int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
    double s = 0.0;
    for (int64_t i = lo; i < hi; ++i) s = s + a[i] * b[i];
    part[t] = s;
    if (t == 0) T = nt;
  }
  double s = 0.0;
  for (int t = 0; t < T; ++t) s += part[t];
  return s;
}

/* Jacobi: d(i,j,k) = sum_l G11(l,j,k) D(l,i)^2 + sum_l G22(i,l,k) D(l,j)^2
 *                  + sum_l G33(i,j,l) D(l,k)^2,  d = h1 d [+ h2 B],
 * then gs ADD and d <- 1/d. */
void or_jacobi_update(const or_sem_t *s, const double *h1, const double *h2,
                      int ifh2, double *dinv) {
  const int lx = s->lx, np = lx * lx * lx;
  const double *D = s->D;
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (int e = 0; e < s->nelv; ++e) {
    const int64_t o = (int64_t)e * np;
    for (int k = 0; k < lx; ++k)
      for (int j = 0; j < lx; ++j)
        for (int i = 0; i < lx; ++i) {
          double d = 0.0;
          for (int l = 0; l < lx; ++l)
            d += s->G11[o + l + lx * (j + lx * k)] * D[l * lx + i] *
                 D[l * lx + i];
          for (int l = 0; l < lx; ++l)
            d += s->G22[o + i + lx * (l + lx * k)] * D[l * lx + j] *
                 D[l * lx + j];
          for (int l = 0; l < lx; ++l)
            d += s->G33[o + i + lx * (j + lx * l)] * D[l * lx + k] *
                 D[l * lx + k];
          int64_t g = o + i + lx * (j + lx * k);
          d = d * h1[g];
          if (ifh2) d = d + h2[g] * s->B[g];
          dinv[g] = d;
        }
  }
  or_gs_op(dinv, s->gs_nb, s->gs_m, s->gs_cnt, s->gs_idx, OR_GS_ADD);
  for (int64_t i = 0; i < s->n; ++i) dinv[i] = 1.0 / dinv[i];
}

/* Jacobi-preconditioned CG, x0 = 0, r = f (f already masked by the caller),
 * stops after the first iteration with sqrt(<r,r>_mult)/sqrt(vol) < abstol. */
void or_cg_solve(const or_sem_t *s, const double *h1, const double *h2,
                 int ifh2, const double *dinv, double *x, const double *f,
                 double abstol, int max_iter, int *iter, double *res_start,
                 double *res_final) {
  const int64_t n = s->n;
  double *r = (double *)malloc(sizeof(double) * n * 4);
  double *p = r + n, *z = r + 2 * n, *w = r + 3 * n;
  const double norm_fac = 1.0 / sqrt(s->volume);
  for (int64_t i = 0; i < n; ++i) {
    x[i] = 0.0;
    r[i] = f[i];
    p[i] = 0.0;
  }
  double rtr = or_glsc3(n, r, s->mult, r);
  double rnorm = sqrt(rtr) * norm_fac;
  *res_start = rnorm;
  *res_final = rnorm;
  *iter = 0;
  if (rnorm == 0.0) {
    free(r);
    return;
  }
  double rtz1 = 1.0, rtz2;
  int it;
  /* OR_PROGRESS set: a line on stderr every 200 iterations (long solves at
   * full size) */
  const int progress = getenv("OR_PROGRESS") != NULL;
  for (it = 1; it <= max_iter; ++it) {
    for (int64_t i = 0; i < n; ++i) z[i] = r[i] * dinv[i];
    rtz2 = rtz1;
    rtz1 = or_glsc3(n, r, s->mult, z);
    double beta = rtz1 / rtz2;
    if (it == 1) beta = 0.0;
    for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    or_ax_helm(s->lx, s->nelv, w, p, s->D, h1, h2, s->G11, s->G12, s->G13,
               s->G22, s->G23, s->G33, s->B, ifh2);
    or_gs_op(w, s->gs_nb, s->gs_m, s->gs_cnt, s->gs_idx, OR_GS_ADD);
    for (int64_t q = 0; q < s->nmask; ++q) w[s->mask[q]] = 0.0;
    double pap = or_glsc3(n, w, s->mult, p);
    double alpha = rtz1 / pap;
    for (int64_t i = 0; i < n; ++i) {
      x[i] = x[i] + alpha * p[i];
      r[i] = r[i] - alpha * w[i];
    }
    rtr = or_glsc3(n, r, s->mult, r);
    rnorm = sqrt(rtr) * norm_fac;
    if (progress && it % 200 == 0)
      fprintf(stderr, "oracle cg: iteration %d, residual %.3e\n", it, rnorm);
    if (rnorm < abstol) break;
  }
  *iter = it > max_iter ? max_iter : it;
  *res_final = rnorm;
  free(r);
}

/* U[0,1) per global point of the box lattice: or_u01(seed, gid(l)) -- the
 * seeded design of the workloads (continuous by construction). */
void or_box_rng_uniform(int nx, int ny, int nz, int lx, uint64_t seed,
                        double *out) {
  const int64_t n = (int64_t)nx * ny * nz * lx * lx * lx;
  int64_t *gid = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
  or_dofmap_gids(nx, ny, nz, lx, gid);
  for (int64_t l = 0; l < n; ++l) out[l] = or_u01(seed, (uint64_t)gid[l]);
  free(gid);
}
