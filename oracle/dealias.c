/*
 * dealias.c -- FP64 CPU restatement of the dealiased (over-integrated)
 * advection operators: the fields are interpolated from the lx GLL points to
 * lxd Gauss-Legendre (GL) points per direction, the products are formed and
 * weighted there, and the result is mapped back with the transpose of the
 * interpolation (a weak-form projection).
 *
 * TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * sources/adjoint/adv_adjoint_dealias.f90 (adv_lin_dealias_t):
 *   init_dealias          :138-213  GL space of lxd points, GLL->GL
 *                                   interpolator, the GL metric = the GLL
 *                                   cofactors interpolated to GL (:155-163)
 *   compute_adjoint       :227-449  u . grad U^T and the weak term
 *                                   int grad v . (U (x) u), CPU branch
 *                                   :346-445 restated per element
 *   compute_linear        :466-653  u' . grad U + U . grad u'
 *   lxd = 3 (p + 1) / 2   advection_adjoint_fctry.f90:66-70
 * and the forward convective term of Neko's adv_dealias_t%compute (Neko is
 * not vendored; restated): f_c -= J^T [ t . (B grad t_c) ] with t = J u on
 * the GL points.  It is tied to the reference's linearised operator by the
 * exact quadratic identity F(U + e u') = F(U) + e L(u'; U) + e^2 F(u')
 * (tests/test_oracle_dealias.py).
 *
 * Operators on the GL points of one element (w3 = GL tensor weights, c_b[a]
 * = the interpolated cofactor d(r_b)/d(x_a) times the Jacobian):
 *   opgrad_a(f) = w3 sum_b c_b[a] d_b f                    (Neko opgrad)
 *   cdtp_a(x)   = sum_b D_b^T (w3 x c_b[a])                 (Neko cdtp with
 *                                                            B / jac = w3)
 * Neko's GL cdtp forms B_GL / jac_GL, which equals w3 to rounding; w3 is used
 * here and on the device.
 *
 * Conventions: Jm[q*lx + i] = l_i(zgl_q) (GLL basis at the GL points), so
 * t = (Jm x Jm x Jm) u and the map back is its transpose; Dg[i*lxd + j] =
 * l'_j(zgl_i) on the GL points; cof[m*ngl + e*lxd^3 + q], m = 3b + a.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* ---- 1-D rules ------------------------------------------------------------*/

/* Gauss-Legendre nodes (ascending) and weights: Newton on P_n from the
 * Chebyshev-like initial guess cos(pi (i + 3/4) / (n + 1/2)). */
void or_gauss_legendre(int n, double *z, double *w) {
  for (int i = 0; i < n; ++i) {
    double x = cos(M_PI * (i + 0.75) / (n + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = x;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      /* P_n = p1, P_{n-1} = p0 */
      dp = n * (x * p1 - p0) / (x * x - 1.0);
      const double dx = p1 / dp;
      x -= dx;
      if (fabs(dx) < 1e-16) break;
    }
    /* recompute P_n' at the converged root */
    double p0 = 1.0, p1 = x;
    for (int k = 2; k <= n; ++k) {
      const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    dp = n * (x * p1 - p0) / (x * x - 1.0);
    z[n - 1 - i] = x;
    w[n - 1 - i] = 2.0 / ((1.0 - x * x) * dp * dp);
  }
}

/* J[q*nf + i] = l_i(zt[q]), the Lagrange basis of the nf nodes zf. */
void or_lagrange_matrix(int nf, const double *zf, int nt, const double *zt,
                        double *J) {
  for (int q = 0; q < nt; ++q)
    for (int i = 0; i < nf; ++i) {
      double v = 1.0;
      for (int k = 0; k < nf; ++k)
        if (k != i) v *= (zt[q] - zf[k]) / (zf[i] - zf[k]);
      J[q * nf + i] = v;
    }
}

/* D[i*n + j] = l_j'(z_i) for the Lagrange basis of the n nodes z:
 * off-diagonal (a_i / a_j) / (z_i - z_j) with a_i = prod_{k!=i} (z_i - z_k),
 * diagonal sum_{k!=i} 1 / (z_i - z_k). */
void or_deriv_matrix(int n, const double *z, double *D) {
  double *a = (double *)malloc(sizeof(double) * n);
  for (int i = 0; i < n; ++i) {
    a[i] = 1.0;
    for (int k = 0; k < n; ++k)
      if (k != i) a[i] *= z[i] - z[k];
  }
  for (int i = 0; i < n; ++i) {
    double d = 0.0;
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      D[i * n + j] = (a[i] / a[j]) / (z[i] - z[j]);
      d += 1.0 / (z[i] - z[j]);
    }
    D[i * n + i] = d;
  }
  free(a);
}

/* ---- element tensor helpers ----------------------------------------------*/

/* out (lxd^3) = (J x J x J) in (lx^3) */
static void interp_up(int lx, int lxd, const double *Jm, const double *in,
                      double *out, double *t1, double *t2) {
  for (int k = 0; k < lx; ++k)
    for (int j = 0; j < lx; ++j)
      for (int q = 0; q < lxd; ++q) {
        double s = 0.0;
        for (int i = 0; i < lx; ++i) s += Jm[q * lx + i] * in[i + lx * (j + lx * k)];
        t1[q + lxd * (j + lx * k)] = s;
      }
  for (int k = 0; k < lx; ++k)
    for (int r = 0; r < lxd; ++r)
      for (int q = 0; q < lxd; ++q) {
        double s = 0.0;
        for (int j = 0; j < lx; ++j) s += Jm[r * lx + j] * t1[q + lxd * (j + lx * k)];
        t2[q + lxd * (r + lxd * k)] = s;
      }
  for (int t = 0; t < lxd; ++t)
    for (int r = 0; r < lxd; ++r)
      for (int q = 0; q < lxd; ++q) {
        double s = 0.0;
        for (int k = 0; k < lx; ++k) s += Jm[t * lx + k] * t2[q + lxd * (r + lxd * k)];
        out[q + lxd * (r + lxd * t)] = s;
      }
}

/* out (lx^3) = (J x J x J)^T in (lxd^3): the map back to GLL */
static void interp_dn(int lx, int lxd, const double *Jm, const double *in,
                      double *out, double *t1, double *t2) {
  for (int t = 0; t < lxd; ++t)
    for (int r = 0; r < lxd; ++r)
      for (int i = 0; i < lx; ++i) {
        double s = 0.0;
        for (int q = 0; q < lxd; ++q) s += Jm[q * lx + i] * in[q + lxd * (r + lxd * t)];
        t2[i + lx * (r + lxd * t)] = s;
      }
  for (int t = 0; t < lxd; ++t)
    for (int j = 0; j < lx; ++j)
      for (int i = 0; i < lx; ++i) {
        double s = 0.0;
        for (int r = 0; r < lxd; ++r) s += Jm[r * lx + j] * t2[i + lx * (r + lxd * t)];
        t1[i + lx * (j + lx * t)] = s;
      }
  for (int k = 0; k < lx; ++k)
    for (int j = 0; j < lx; ++j)
      for (int i = 0; i < lx; ++i) {
        double s = 0.0;
        for (int t = 0; t < lxd; ++t) s += Jm[t * lx + k] * t1[i + lx * (j + lx * t)];
        out[i + lx * (j + lx * k)] = s;
      }
}

/* r/s/t derivatives on the GL points */
static void grad_ref(int n, const double *Dg, const double *f, double *fr,
                     double *fs, double *ft) {
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        double a = 0, b = 0, c = 0;
        for (int m = 0; m < n; ++m) {
          a += Dg[i * n + m] * f[m + n * (j + n * k)];
          b += Dg[j * n + m] * f[i + n * (m + n * k)];
          c += Dg[k * n + m] * f[i + n * (j + n * m)];
        }
        const int p = i + n * (j + n * k);
        fr[p] = a;
        fs[p] = b;
        ft[p] = c;
      }
}

typedef struct {
  int lx, lxd, npg;
  int64_t ngl;
  const double *Jm, *Dg, *wg, *cof;
  /* per-thread scratch */
  double *t1, *t2, *fr, *fs, *ft, *w3;
} gl_t;

static double w3_at(const gl_t *g, int p) {
  const int n = g->lxd;
  return g->wg[p % n] * g->wg[(p / n) % n] * g->wg[p / (n * n)];
}

static double cof_at(const gl_t *g, int e, int m, int p) {
  return g->cof[(int64_t)m * g->ngl + (int64_t)e * g->npg + p];
}

/* opgrad of f (GL) on element e: ox, oy, oz (Neko opgrad on coef_GL) */
static void opgrad_gl(const gl_t *g, int e, const double *f, double *o[3]) {
  grad_ref(g->lxd, g->Dg, f, g->fr, g->fs, g->ft);
  for (int p = 0; p < g->npg; ++p) {
    const double w3 = w3_at(g, p);
    for (int a = 0; a < 3; ++a)
      o[a][p] = w3 * (cof_at(g, e, a, p) * g->fr[p] + cof_at(g, e, 3 + a, p) * g->fs[p] +
                      cof_at(g, e, 6 + a, p) * g->ft[p]);
  }
}

/* Neko cdtp on coef_GL with the cofactor column a: out = sum_b D_b^T (w3 x c_b[a]) */
static void cdtp_gl(const gl_t *g, int e, const double *x, int a, double *out) {
  const int n = g->lxd, np = g->npg;
  for (int p = 0; p < np; ++p) {
    const double wx = w3_at(g, p) * x[p];
    g->fr[p] = wx * cof_at(g, e, a, p);
    g->fs[p] = wx * cof_at(g, e, 3 + a, p);
    g->ft[p] = wx * cof_at(g, e, 6 + a, p);
  }
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int m = 0; m < n; ++m) s += g->Dg[m * n + i] * g->fr[m + n * (j + n * k)];
        double s2 = 0.0;
        for (int m = 0; m < n; ++m) s2 += g->Dg[m * n + j] * g->fs[i + n * (m + n * k)];
        double s3 = 0.0;
        for (int m = 0; m < n; ++m) s3 += g->Dg[m * n + k] * g->ft[i + n * (j + n * m)];
        out[i + n * (j + n * k)] = s + s2 + s3;
      }
}

static void gl_scratch(gl_t *g, double **blk) {
  const int np = g->npg;
  *blk = (double *)malloc(sizeof(double) * (size_t)np * 6);
  g->t1 = *blk;
  g->t2 = g->t1 + np;
  g->fr = g->t2 + np;
  g->fs = g->fr + np;
  g->ft = g->fs + np;
  g->w3 = g->ft + np;
}

/* ---- setup: the GL metric ------------------------------------------------*/

void or_dealias_metric(int lx, int lxd, int nelv, const double *Jm,
                       const double *cof_gll /* [9][nelv*lx^3] */,
                       double *cof /* [9][nelv*lxd^3] */) {
  const int np = lx * lx * lx, npg = lxd * lxd * lxd;
  const int64_t n = (int64_t)nelv * np, ngl = (int64_t)nelv * npg;
#pragma omp parallel num_threads(or_get_threads())
  {
    double *t = (double *)malloc(sizeof(double) * (size_t)npg * 2);
#pragma omp for schedule(static)
    for (int e = 0; e < nelv; ++e)
      for (int m = 0; m < 9; ++m)
        interp_up(lx, lxd, Jm, cof_gll + m * n + (int64_t)e * np,
                  cof + m * ngl + (int64_t)e * npg, t, t + npg);
    free(t);
  }
}

/* ---- the three operators -------------------------------------------------*/

#define OR_GL_SETUP                                                          \
  const int np = lx * lx * lx, npg = lxd * lxd * lxd;                         \
  const int64_t ngl = (int64_t)nelv * npg;                                    \
  (void)ngl

/* forward (Neko adv_dealias): f_c -= J^T [ sum_j t_j opgrad_j(t_c) ] */
void or_dealias_advection(int lx, int lxd, int nelv, const double *Jm,
                          const double *Dg, const double *wg,
                          const double *cof, const double *u, const double *v,
                          const double *w, double *fx, double *fy, double *fz) {
  OR_GL_SETUP;
#pragma omp parallel num_threads(or_get_threads())
  {
    gl_t g = {lx, lxd, npg, ngl, Jm, Dg, wg, cof, 0, 0, 0, 0, 0, 0};
    double *blk;
    gl_scratch(&g, &blk);
    double *buf = (double *)malloc(sizeof(double) * (size_t)npg * 8);
    double *t[3] = {buf, buf + npg, buf + 2 * npg};
    double *o[3] = {buf + 3 * npg, buf + 4 * npg, buf + 5 * npg};
    double *tbf = buf + 6 * npg, *tmp = buf + 7 * npg;
    const double *src[3] = {u, v, w};
    double *dst[3] = {fx, fy, fz};
#pragma omp for schedule(static)
    for (int e = 0; e < nelv; ++e) {
      const int64_t off = (int64_t)e * np;
      for (int a = 0; a < 3; ++a) interp_up(lx, lxd, Jm, src[a] + off, t[a], g.t1, g.t2);
      for (int c = 0; c < 3; ++c) {
        opgrad_gl(&g, e, t[c], o);
        for (int p = 0; p < npg; ++p)
          tbf[p] = o[0][p] * t[0][p] + o[1][p] * t[1][p] + o[2][p] * t[2][p];
        interp_dn(lx, lxd, Jm, tbf, tmp, g.t1, g.t2);
        for (int p = 0; p < np; ++p) dst[c][off + p] -= tmp[p];
      }
    }
    free(buf);
    free(blk);
  }
}

/* linearised (compute_linear_advection_dealias :466-653, device branch
 * :500-545): f_c -= J^T [ sum_j up_j opgrad_j(U_c) ];
 *            f_c -= J^T [ sum_j U_j opgrad_j(up_c) ] */
void or_dealias_linear(int lx, int lxd, int nelv, const double *Jm,
                       const double *Dg, const double *wg, const double *cof,
                       const double *up, const double *vp, const double *wp,
                       const double *ub, const double *vb, const double *wb,
                       double *fx, double *fy, double *fz) {
  OR_GL_SETUP;
#pragma omp parallel num_threads(or_get_threads())
  {
    gl_t g = {lx, lxd, npg, ngl, Jm, Dg, wg, cof, 0, 0, 0, 0, 0, 0};
    double *blk;
    gl_scratch(&g, &blk);
    double *buf = (double *)malloc(sizeof(double) * (size_t)npg * 11);
    double *t[3] = {buf, buf + npg, buf + 2 * npg};
    double *tb[3] = {buf + 3 * npg, buf + 4 * npg, buf + 5 * npg};
    double *o[3] = {buf + 6 * npg, buf + 7 * npg, buf + 8 * npg};
    double *tbf = buf + 9 * npg, *tmp = buf + 10 * npg;
    const double *sp[3] = {up, vp, wp}, *sb[3] = {ub, vb, wb};
    double *dst[3] = {fx, fy, fz};
#pragma omp for schedule(static)
    for (int e = 0; e < nelv; ++e) {
      const int64_t off = (int64_t)e * np;
      for (int a = 0; a < 3; ++a) {
        interp_up(lx, lxd, Jm, sb[a] + off, tb[a], g.t1, g.t2);
        interp_up(lx, lxd, Jm, sp[a] + off, t[a], g.t1, g.t2);
      }
      for (int c = 0; c < 3; ++c) { /* u' . grad U_c */
        opgrad_gl(&g, e, tb[c], o);
        for (int p = 0; p < npg; ++p)
          tbf[p] = o[0][p] * t[0][p] + o[1][p] * t[1][p] + o[2][p] * t[2][p];
        interp_dn(lx, lxd, Jm, tbf, tmp, g.t1, g.t2);
        for (int p = 0; p < np; ++p) dst[c][off + p] -= tmp[p];
      }
      for (int c = 0; c < 3; ++c) { /* U . grad u'_c */
        opgrad_gl(&g, e, t[c], o);
        for (int p = 0; p < npg; ++p)
          tbf[p] = o[0][p] * tb[0][p] + o[1][p] * tb[1][p] + o[2][p] * tb[2][p];
        interp_dn(lx, lxd, Jm, tbf, tmp, g.t1, g.t2);
        for (int p = 0; p < np; ++p) dst[c][off + p] -= tmp[p];
      }
    }
    free(buf);
    free(blk);
  }
}

/* adjoint (compute_adjoint_advection_dealias :227-449, per element as the
 * CPU branch :346-445):
 *   tf_c = sum_d a_d opgrad_c(U_d);  f_c -= J^T tf_c          (:366-381)
 *   for each component c:  f_c -= J^T [ cdtp_x(a_c U) + cdtp_y(a_c V)
 *                                      + cdtp_z(a_c W) ]        (:383-441) */
void or_dealias_adjoint(int lx, int lxd, int nelv, const double *Jm,
                        const double *Dg, const double *wg, const double *cof,
                        const double *ua, const double *va, const double *wa,
                        const double *ub, const double *vb, const double *wb,
                        double *fx, double *fy, double *fz) {
  OR_GL_SETUP;
#pragma omp parallel num_threads(or_get_threads())
  {
    gl_t g = {lx, lxd, npg, ngl, Jm, Dg, wg, cof, 0, 0, 0, 0, 0, 0};
    double *blk;
    gl_scratch(&g, &blk);
    double *buf = (double *)malloc(sizeof(double) * (size_t)npg * 23);
    double *t[3], *tb[3], *du[3][3], *tf[3], *pr[3], *tmp = buf + 22 * npg;
    for (int a = 0; a < 3; ++a) {
      t[a] = buf + a * npg;
      tb[a] = buf + (3 + a) * npg;
      tf[a] = buf + (15 + a) * npg;
      pr[a] = buf + (18 + a) * npg;
      for (int b = 0; b < 3; ++b) du[a][b] = buf + (6 + 3 * a + b) * npg;
    }
    double *sum = buf + 21 * npg;
    const double *sa[3] = {ua, va, wa}, *sb[3] = {ub, vb, wb};
    double *dst[3] = {fx, fy, fz};
#pragma omp for schedule(static)
    for (int e = 0; e < nelv; ++e) {
      const int64_t off = (int64_t)e * np;
      for (int a = 0; a < 3; ++a) {
        interp_up(lx, lxd, Jm, sb[a] + off, tb[a], g.t1, g.t2);
        interp_up(lx, lxd, Jm, sa[a] + off, t[a], g.t1, g.t2);
      }
      /* du[d][c] = opgrad_c(U_d) */
      for (int d = 0; d < 3; ++d) opgrad_gl(&g, e, tb[d], du[d]);
      for (int c = 0; c < 3; ++c) {
        for (int p = 0; p < npg; ++p)
          tf[c][p] = t[0][p] * du[0][c][p] + t[1][p] * du[1][c][p] + t[2][p] * du[2][c][p];
        interp_dn(lx, lxd, Jm, tf[c], tmp, g.t1, g.t2);
        for (int p = 0; p < np; ++p) dst[c][off + p] -= tmp[p];
      }
      for (int c = 0; c < 3; ++c) {
        for (int d = 0; d < 3; ++d) {
          for (int p = 0; p < npg; ++p) sum[p] = t[c][p] * tb[d][p];
          cdtp_gl(&g, e, sum, d, pr[d]);
        }
        for (int p = 0; p < npg; ++p) sum[p] = pr[0][p] + pr[1][p] + pr[2][p];
        interp_dn(lx, lxd, Jm, sum, tmp, g.t1, g.t2);
        for (int p = 0; p < np; ++p) dst[c][off + p] -= tmp[p];
      }
    }
    free(buf);
    free(blk);
  }
}
