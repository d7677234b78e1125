/*
 * mma.c -- FP64 CPU restatement of neko-top's MMA (method of moving
 * asymptotes, Svanberg): front end sources/mma/mma.f90:132-324, sub-problem
 * generation, dual-primal interior-point sub-solve and KKT residual
 * sources/mma/bcknd/cpu/mma_cpu.f90:8-641.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle).  The steps, their order and the
 * reference's default-real (single precision) literals are kept so results
 * can be compared at FP64 round-off: F0_01 = real(0.01), F0_1 = real(0.1),
 * F0_9 = real(0.9), F1_01 = real(1.01).  One rank: the reference's
 * MPI_Allreduce calls are identities.  dfdx and pij/qij are m x n
 * column-major (element (i,j) at [j*m + i]), as the Fortran arrays.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define F0_01 ((double)0.01f)
#define F0_1 ((double)0.1f)
#define F0_9 ((double)0.9f)
#define F1_01 ((double)1.01f)

struct or_mma_s {
  int n, m, max_iter;
  double a0, asyinit, asyincr, asydecr, epsimin, residumax, residunorm, z,
      zeta;
  double *xold1, *xold2, *low, *upp, *alpha, *beta, *xmin, *xmax, *p0j, *q0j,
      *xsi, *eta;
  double *pij, *qij;
  double *a, *c, *d, *bi, *y, *lambda, *s, *mu;
};

static double *dup(const double *v, int n) {
  double *p = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
  if (v) memcpy(p, v, sizeof(double) * n);
  else memset(p, 0, sizeof(double) * (n > 0 ? n : 1));
  return p;
}

/* mma.f90:132-184 (mma_init_attributes); epsimin <= 0 selects the default
 * 1e-9 sqrt(m + n). */
or_mma_t *or_mma_create(int n, int m, const double *x, double a0,
                        const double *a, const double *c, const double *d,
                        const double *xmin, const double *xmax, int max_iter,
                        double epsimin, double asyinit, double asyincr,
                        double asydecr) {
  or_mma_t *t = (or_mma_t *)calloc(1, sizeof(or_mma_t));
  t->n = n;
  t->m = m;
  t->max_iter = max_iter;
  t->a0 = a0;
  t->asyinit = asyinit;
  t->asyincr = asyincr;
  t->asydecr = asydecr;
  t->epsimin = epsimin > 0 ? epsimin : 1.0e-9 * sqrt((double)(m + n));
  t->xold1 = dup(x, n);
  t->xold2 = dup(x, n);
  t->xmin = dup(xmin, n);
  t->xmax = dup(xmax, n);
  double lo = x[0], hi = x[0];
  for (int j = 1; j < n; ++j) {
    if (x[j] < lo) lo = x[j];
    if (x[j] > hi) hi = x[j];
  }
  t->low = dup(NULL, n);
  t->upp = dup(NULL, n);
  for (int j = 0; j < n; ++j) {
    t->low[j] = lo;
    t->upp[j] = hi;
  }
  t->alpha = dup(NULL, n);
  t->beta = dup(NULL, n);
  t->p0j = dup(NULL, n);
  t->q0j = dup(NULL, n);
  t->xsi = dup(NULL, n);
  t->eta = dup(NULL, n);
  t->pij = dup(NULL, m * n);
  t->qij = dup(NULL, m * n);
  t->a = dup(a, m);
  t->c = dup(c, m);
  t->d = dup(d, m);
  t->bi = dup(NULL, m);
  t->y = dup(NULL, m);
  t->lambda = dup(NULL, m);
  t->s = dup(NULL, m);
  t->mu = dup(NULL, m);
  t->residumax = HUGE_VAL;
  t->residunorm = HUGE_VAL;
  return t;
}

void or_mma_free(or_mma_t *t) {
  if (!t) return;
  double *v[] = {t->xold1, t->xold2, t->low, t->upp, t->alpha, t->beta,
                 t->xmin, t->xmax, t->p0j, t->q0j, t->xsi, t->eta, t->pij,
                 t->qij, t->a, t->c, t->d, t->bi, t->y, t->lambda, t->s, t->mu};
  for (size_t q = 0; q < sizeof(v) / sizeof(v[0]); ++q) free(v[q]);
  free(t);
}

double or_mma_residumax(const or_mma_t *t) { return t->residumax; }
double or_mma_residunorm(const or_mma_t *t) { return t->residunorm; }

/* Copy of the sub-problem state for parity tests: low, upp, alpha, beta,
 * p0j, q0j (n each), pij, qij (m*n each), bi (m). */
void or_mma_get_subproblem(const or_mma_t *t, double *low, double *upp,
                           double *alpha, double *beta, double *p0j,
                           double *q0j, double *pij, double *qij, double *bi) {
  const size_t n = t->n, mn = (size_t)t->m * t->n;
  memcpy(low, t->low, n * 8);
  memcpy(upp, t->upp, n * 8);
  memcpy(alpha, t->alpha, n * 8);
  memcpy(beta, t->beta, n * 8);
  memcpy(p0j, t->p0j, n * 8);
  memcpy(q0j, t->q0j, n * 8);
  memcpy(pij, t->pij, mn * 8);
  memcpy(qij, t->qij, mn * 8);
  memcpy(bi, t->bi, (size_t)t->m * 8);
}

/* Duals after a sub-solve: y, lambda, s, mu (m each), xsi, eta (n each),
 * zs = {z, zeta}. */
void or_mma_get_duals(const or_mma_t *t, double *y, double *lambda, double *s,
                      double *mu, double *xsi, double *eta, double *zs) {
  memcpy(y, t->y, (size_t)t->m * 8);
  memcpy(lambda, t->lambda, (size_t)t->m * 8);
  memcpy(s, t->s, (size_t)t->m * 8);
  memcpy(mu, t->mu, (size_t)t->m * 8);
  memcpy(xsi, t->xsi, (size_t)t->n * 8);
  memcpy(eta, t->eta, (size_t)t->n * 8);
  zs[0] = t->z;
  zs[1] = t->zeta;
}

/* mma_cpu.f90:8-184 */
static void gensub(or_mma_t *t, int iter, const double *x, const double *df0dx,
                   const double *fval, const double *dfdx) {
  const int n = t->n, m = t->m;
  if (iter < 3) {
    for (int j = 0; j < n; ++j) {
      t->low[j] = x[j] - t->asyinit * (t->xmax[j] - t->xmin[j]);
      t->upp[j] = x[j] + t->asyinit * (t->xmax[j] - t->xmin[j]);
    }
  } else {
    for (int j = 0; j < n; ++j) {
      const double osc = (x[j] - t->xold1[j]) * (t->xold1[j] - t->xold2[j]);
      if (osc < 0) {
        t->low[j] = x[j] - t->asydecr * (t->xold1[j] - t->low[j]);
        t->upp[j] = x[j] + t->asydecr * (t->upp[j] - t->xold1[j]);
      } else if (osc > 0) {
        t->low[j] = x[j] - t->asyincr * (t->xold1[j] - t->low[j]);
        t->upp[j] = x[j] + t->asyincr * (t->upp[j] - t->xold1[j]);
      } else {
        t->low[j] = x[j] - (t->xold1[j] - t->low[j]);
        t->upp[j] = x[j] + (t->upp[j] - t->xold1[j]);
      }
      const double w = t->xmax[j] - t->xmin[j];
      t->low[j] = fmax(t->low[j], x[j] - 10.0 * w);
      t->low[j] = fmin(t->low[j], x[j] - F0_01 * w);
      t->upp[j] = fmin(t->upp[j], x[j] + 10.0 * w);
      t->upp[j] = fmax(t->upp[j], x[j] + F0_01 * w);
    }
  }
  for (int j = 0; j < n; ++j) {
    const double w = t->xmax[j] - t->xmin[j];
    double al = t->low[j] + 0.1 * (x[j] - t->low[j]);
    double a2 = x[j] - 0.5 * w;
    t->alpha[j] = fmax(fmax(t->xmin[j], al), a2);
    double be = t->upp[j] - F0_1 * (t->upp[j] - x[j]);
    double b2 = x[j] + 0.5 * w;
    t->beta[j] = fmin(fmin(t->xmax[j], be), b2);
    const double reg = 0.00001 / fmax(0.00001, w);
    const double du = (t->upp[j] - x[j]) * (t->upp[j] - x[j]);
    const double dl = (x[j] - t->low[j]) * (x[j] - t->low[j]);
    t->p0j[j] = du * (1.001 * fmax(df0dx[j], 0.0) + 0.001 * fmax(-df0dx[j], 0.0) + reg);
    t->q0j[j] = dl * (0.001 * fmax(df0dx[j], 0.0) + 1.001 * fmax(-df0dx[j], 0.0) + reg);
    for (int i = 0; i < m; ++i) {
      const double g = dfdx[(size_t)j * m + i];
      t->pij[(size_t)j * m + i] = du * (1.001 * fmax(g, 0.0) + 0.001 * fmax(-g, 0.0) + reg);
      t->qij[(size_t)j * m + i] = dl * (0.001 * fmax(g, 0.0) + 1.001 * fmax(-g, 0.0) + reg);
    }
  }
  for (int i = 0; i < m; ++i) {
    double b = 0.0;
    for (int j = 0; j < n; ++j)
      b = b + t->pij[(size_t)j * m + i] / (t->upp[j] - x[j]) +
          t->qij[(size_t)j * m + i] / (x[j] - t->low[j]);
    t->bi[i] = b - fval[i];
  }
}

/* DGESV semantics: LU with row partial pivoting, one right-hand side,
 * column-major A (N x N).  Returns 0, or k > 0 when U(k,k) is exactly 0. */
int or_dgesv(int N, double *A, double *b) {
  int info = 0;
  for (int k = 0; k < N; ++k) {
    int p = k;
    double big = fabs(A[k + k * N]);
    for (int r = k + 1; r < N; ++r)
      if (fabs(A[r + k * N]) > big) {
        big = fabs(A[r + k * N]);
        p = r;
      }
    if (p != k) {
      for (int c = 0; c < N; ++c) {
        double tmp = A[k + c * N];
        A[k + c * N] = A[p + c * N];
        A[p + c * N] = tmp;
      }
      double tb = b[k];
      b[k] = b[p];
      b[p] = tb;
    }
    if (A[k + k * N] == 0.0) {
      if (!info) info = k + 1;
      continue;
    }
    for (int r = k + 1; r < N; ++r) {
      const double l = A[r + k * N] / A[k + k * N];
      A[r + k * N] = l;
      for (int c = k + 1; c < N; ++c) A[r + c * N] -= l * A[k + c * N];
      b[r] -= l * b[k];
    }
  }
  if (info) return info;
  for (int k = N - 1; k >= 0; --k) {
    double s = b[k];
    for (int c = k + 1; c < N; ++c) s -= A[k + c * N] * b[c];
    b[k] = s / A[k + k * N];
  }
  return 0;
}

typedef struct {
  double *rex, *rexsi, *reeta, *rey, *relambda, *remu, *res;
  double rez, rezeta, max, norm;
} resid_t;

/* Residuals of the relaxed KKT system, mma_cpu.f90:251-315 / :495-539. */
static void residuals(const or_mma_t *t, const double *x, const double *y,
                      double z, const double *lambda, const double *xsi,
                      const double *eta, const double *mu, double zeta,
                      const double *s, double epsi, resid_t *r) {
  const int n = t->n, m = t->m;
  double xs = 0.0, mx = 0.0, small = 0.0;
  for (int j = 0; j < n; ++j) {
    double pl = t->p0j[j], ql = t->q0j[j];
    for (int i = 0; i < m; ++i) {
      pl += t->pij[(size_t)j * m + i] * lambda[i];
      ql += t->qij[(size_t)j * m + i] * lambda[i];
    }
    const double ux = t->upp[j] - x[j], xl = x[j] - t->low[j];
    r->rex[j] = (pl / (ux * ux) - ql / (xl * xl)) - xsi[j] + eta[j];
    r->rexsi[j] = xsi[j] * (x[j] - t->alpha[j]) - epsi;
    r->reeta[j] = eta[j] * (t->beta[j] - x[j]) - epsi;
    xs += r->rex[j] * r->rex[j];
  }
  double xs2 = 0.0, xs3 = 0.0;
  for (int j = 0; j < n; ++j) {
    xs2 += r->rexsi[j] * r->rexsi[j];
    xs3 += r->reeta[j] * r->reeta[j];
  }
  for (int i = 0; i < m; ++i) {
    r->rey[i] = t->c[i] + t->d[i] * y[i] - lambda[i] - mu[i];
    double g = 0.0;
    for (int j = 0; j < n; ++j)
      g = g + t->pij[(size_t)j * m + i] / (t->upp[j] - x[j]) +
          t->qij[(size_t)j * m + i] / (x[j] - t->low[j]);
    r->relambda[i] = g - t->a[i] * z - y[i] + s[i] - t->bi[i];
    r->remu[i] = mu[i] * y[i] - epsi;
    r->res[i] = lambda[i] * s[i] - epsi;
  }
  double la = 0.0;
  for (int i = 0; i < m; ++i) la += lambda[i] * t->a[i];
  r->rez = t->a0 - zeta - la;
  r->rezeta = zeta * z - epsi;
  for (int j = 0; j < n; ++j) {
    mx = fmax(mx, fabs(r->rex[j]));
    mx = fmax(mx, fabs(r->rexsi[j]));
    mx = fmax(mx, fabs(r->reeta[j]));
  }
  for (int i = 0; i < m; ++i) {
    mx = fmax(mx, fabs(r->rey[i]));
    mx = fmax(mx, fabs(r->relambda[i]));
    mx = fmax(mx, fabs(r->remu[i]));
    mx = fmax(mx, fabs(r->res[i]));
    small += r->rey[i] * r->rey[i] + r->relambda[i] * r->relambda[i] +
             r->remu[i] * r->remu[i] + r->res[i] * r->res[i];
  }
  mx = fmax(mx, fabs(r->rez));
  mx = fmax(mx, fabs(r->rezeta));
  small += r->rez * r->rez + r->rezeta * r->rezeta;
  r->max = mx;
  r->norm = sqrt(small + (xs + xs2 + xs3));
}

/* mma_cpu.f90:186-579.  Returns 0 or the DGESV failure code. */
static int subsolve(or_mma_t *t, double *designx) {
  const int n = t->n, m = t->m;
  double *nb = (double *)calloc((size_t)n * 16 + 8, sizeof(double));
  double *x = nb, *xsi = nb + n, *eta = nb + 2 * n, *dx = nb + 3 * n,
         *dxsi = nb + 4 * n, *deta = nb + 5 * n, *xold = nb + 6 * n,
         *xsiold = nb + 7 * n, *etaold = nb + 8 * n, *delx = nb + 9 * n,
         *diagx = nb + 10 * n;
  resid_t R;
  R.rex = nb + 11 * n;
  R.rexsi = nb + 12 * n;
  R.reeta = nb + 13 * n;
  double *mb = (double *)calloc((size_t)m * 24 + (m + 1) * (m + 1) + 8, sizeof(double));
  double *y = mb, *lambda = mb + m, *s = mb + 2 * m, *mu = mb + 3 * m,
         *dy = mb + 4 * m, *dlambda = mb + 5 * m, *ds = mb + 6 * m,
         *dmu = mb + 7 * m, *yold = mb + 8 * m, *lambdaold = mb + 9 * m,
         *sold = mb + 10 * m, *muold = mb + 11 * m, *dely = mb + 12 * m,
         *dellambda = mb + 13 * m, *bb = mb + 14 * m; /* m+1 */
  R.rey = mb + 16 * m;
  R.relambda = mb + 17 * m;
  R.remu = mb + 18 * m;
  R.res = mb + 19 * m;
  double *AA = mb + 24 * m;
  int rc = 0;

  double epsi = 1.0;
  double z = 1.0, zeta = 1.0;
  for (int j = 0; j < n; ++j) {
    x[j] = 0.5 * (t->alpha[j] + t->beta[j]);
    xsi[j] = fmax(1.0, 1.0 / (x[j] - t->alpha[j]));
    eta[j] = fmax(1.0, 1.0 / (t->beta[j] - x[j]));
  }
  for (int i = 0; i < m; ++i) {
    y[i] = 1.0;
    lambda[i] = 1.0;
    s[i] = 1.0;
    mu[i] = fmax(1.0, 0.5 * t->c[i]);
  }
  while (epsi > F0_9 * t->epsimin) {
    residuals(t, x, y, z, lambda, xsi, eta, mu, zeta, s, epsi, &R);
    double residumax = R.max, residunorm = R.norm;
    for (int it = 1; it <= t->max_iter; ++it) {
      if (residumax < epsi) break;
      for (int j = 0; j < n; ++j) {
        const double ux = t->upp[j] - x[j], xl = x[j] - t->low[j];
        double a = 0.0, pl = t->p0j[j], ql = t->q0j[j];
        for (int i = 0; i < m; ++i) {
          a = a + t->pij[(size_t)j * m + i] * lambda[i] / (ux * ux) -
              t->qij[(size_t)j * m + i] * lambda[i] / (xl * xl);
          pl += t->pij[(size_t)j * m + i] * lambda[i];
          ql += t->qij[(size_t)j * m + i] * lambda[i];
        }
        delx[j] = a + t->p0j[j] / (ux * ux) - t->q0j[j] / (xl * xl) -
                  epsi / (x[j] - t->alpha[j]) + epsi / (t->beta[j] - x[j]);
        const double dg = pl / (ux * ux * ux) + ql / (xl * xl * xl);
        diagx[j] = 2.0 * dg + xsi[j] / (x[j] - t->alpha[j]) +
                   eta[j] / (t->beta[j] - x[j]);
      }
      double la = 0.0;
      for (int i = 0; i < m; ++i) la += lambda[i] * t->a[i];
      const double delz = t->a0 - la - epsi / z;
      for (int i = 0; i < m; ++i) {
        dely[i] = t->c[i] + t->d[i] * y[i] - lambda[i] - epsi / y[i];
        double g = 0.0, b = 0.0;
        for (int j = 0; j < n; ++j) {
          const double ux = t->upp[j] - x[j], xl = x[j] - t->low[j];
          g = g + t->pij[(size_t)j * m + i] / ux + t->qij[(size_t)j * m + i] / xl;
          const double GG = t->pij[(size_t)j * m + i] / (ux * ux) -
                            t->qij[(size_t)j * m + i] / (xl * xl);
          b = b + GG * (delx[j] / diagx[j]);
        }
        dellambda[i] = g - t->a[i] * z - y[i] - t->bi[i] + epsi / lambda[i];
        bb[i] = dellambda[i] + dely[i] / (t->d[i] + mu[i] / y[i]) - b;
      }
      bb[m] = delz;
      const int M1 = m + 1;
      for (int q = 0; q < M1 * M1; ++q) AA[q] = 0.0;
      for (int i = 0; i < m; ++i)
        for (int k = 0; k < m; ++k) {
          double acc = 0.0;
          for (int j = 0; j < n; ++j) {
            const double ux = t->upp[j] - x[j], xl = x[j] - t->low[j];
            const double Gi = t->pij[(size_t)j * m + i] / (ux * ux) -
                              t->qij[(size_t)j * m + i] / (xl * xl);
            const double Gk = t->pij[(size_t)j * m + k] / (ux * ux) -
                              t->qij[(size_t)j * m + k] / (xl * xl);
            acc = acc + Gi * (1.0 / diagx[j]) * Gk;
          }
          AA[i + k * M1] = acc;
        }
      for (int i = 0; i < m; ++i) {
        AA[i + i * M1] += s[i] / lambda[i] + 1.0 / (t->d[i] + mu[i] / y[i]);
        AA[i + m * M1] = t->a[i];
        AA[m + i * M1] = t->a[i];
      }
      AA[m + m * M1] = -zeta / z;
      int info = or_dgesv(M1, AA, bb);
      if (info) {
        rc = info;
        goto done;
      }
      for (int i = 0; i < m; ++i) dlambda[i] = bb[i];
      const double dz = bb[m];
      double steg = 1.0;
      for (int j = 0; j < n; ++j) {
        const double ux = t->upp[j] - x[j], xl = x[j] - t->low[j];
        double gl = 0.0;
        for (int i = 0; i < m; ++i)
          gl += (t->pij[(size_t)j * m + i] / (ux * ux) -
                 t->qij[(size_t)j * m + i] / (xl * xl)) * dlambda[i];
        dx[j] = -delx[j] / diagx[j] - gl / diagx[j];
        dxsi[j] = -xsi[j] + (epsi - dx[j] * xsi[j]) / (x[j] - t->alpha[j]);
        deta[j] = -eta[j] + (epsi + dx[j] * eta[j]) / (t->beta[j] - x[j]);
      }
      for (int i = 0; i < m; ++i) {
        dy[i] = (-dely[i] + dlambda[i]) / (t->d[i] + mu[i] / y[i]);
        dmu[i] = -mu[i] + (epsi - mu[i] * dy[i]) / y[i];
        ds[i] = -s[i] + (epsi - dlambda[i] * s[i]) / lambda[i];
      }
      const double dzeta = -zeta + (epsi - zeta * dz) / z;
      /* steg = 1 / max(1, -1.01 dxx/xx, -1.01 dx/(x-alpha), 1.01 dx/(beta-x)) */
      for (int i = 0; i < m; ++i) {
        steg = fmax(steg, -(F1_01 * dy[i] / y[i]));
        steg = fmax(steg, -(F1_01 * dlambda[i] / lambda[i]));
        steg = fmax(steg, -(F1_01 * dmu[i] / mu[i]));
        steg = fmax(steg, -(F1_01 * ds[i] / s[i]));
      }
      steg = fmax(steg, -(F1_01 * dz / z));
      steg = fmax(steg, -(F1_01 * dzeta / zeta));
      for (int j = 0; j < n; ++j) {
        steg = fmax(steg, -(F1_01 * dxsi[j] / xsi[j]));
        steg = fmax(steg, -(F1_01 * deta[j] / eta[j]));
        steg = fmax(steg, -(F1_01 * dx[j] / (x[j] - t->alpha[j])));
        steg = fmax(steg, F1_01 * dx[j] / (t->beta[j] - x[j]));
      }
      steg = 1.0 / steg;
      memcpy(xold, x, (size_t)n * 8);
      memcpy(xsiold, xsi, (size_t)n * 8);
      memcpy(etaold, eta, (size_t)n * 8);
      memcpy(yold, y, (size_t)m * 8);
      memcpy(lambdaold, lambda, (size_t)m * 8);
      memcpy(sold, s, (size_t)m * 8);
      memcpy(muold, mu, (size_t)m * 8);
      const double zold = z, zetaold = zeta;
      double newresidu = 2.0 * residunorm;
      int itto = 0;
      while (newresidu > residunorm && itto < 50) {
        ++itto;
        for (int j = 0; j < n; ++j) {
          x[j] = xold[j] + steg * dx[j];
          xsi[j] = xsiold[j] + steg * dxsi[j];
          eta[j] = etaold[j] + steg * deta[j];
        }
        for (int i = 0; i < m; ++i) {
          y[i] = yold[i] + steg * dy[i];
          lambda[i] = lambdaold[i] + steg * dlambda[i];
          mu[i] = muold[i] + steg * dmu[i];
          s[i] = sold[i] + steg * ds[i];
        }
        z = zold + steg * dz;
        zeta = zetaold + steg * dzeta;
        residuals(t, x, y, z, lambda, xsi, eta, mu, zeta, s, epsi, &R);
        newresidu = R.norm;
        steg = steg / 2.0;
      }
      residunorm = newresidu;
      residumax = R.max;
      steg = 2.0 * steg;
    }
    epsi = 0.1 * epsi;
  }
  memcpy(t->xold2, t->xold1, (size_t)n * 8);
  memcpy(t->xold1, designx, (size_t)n * 8);
  memcpy(designx, x, (size_t)n * 8);
  memcpy(t->y, y, (size_t)m * 8);
  memcpy(t->lambda, lambda, (size_t)m * 8);
  memcpy(t->s, s, (size_t)m * 8);
  memcpy(t->mu, mu, (size_t)m * 8);
  memcpy(t->xsi, xsi, (size_t)n * 8);
  memcpy(t->eta, eta, (size_t)n * 8);
  t->z = z;
  t->zeta = zeta;
done:
  free(nb);
  free(mb);
  return rc;
}

/* mma.f90:297-324 (mma_update_cpu): gensub then subsolve; x updated in
 * place.  Returns 0 or the DGESV failure code (the reference stops). */
int or_mma_update(or_mma_t *t, int iter, double *x, const double *df0dx,
                  const double *fval, const double *dfdx) {
  gensub(t, iter, x, df0dx, fval, dfdx);
  return subsolve(t, x);
}

/* mma_cpu.f90:581-641 */
void or_mma_kkt(or_mma_t *t, const double *x, const double *df0dx,
                const double *fval, const double *dfdx) {
  const int n = t->n, m = t->m;
  double mx = 0.0, xs = 0.0, small = 0.0;
  for (int j = 0; j < n; ++j) {
    double g = 0.0;
    for (int i = 0; i < m; ++i) g += dfdx[(size_t)j * m + i] * t->lambda[i];
    const double rex = df0dx[j] + g - t->xsi[j] + t->eta[j];
    const double rexsi = t->xsi[j] * (x[j] - t->xmin[j]);
    const double reeta = t->eta[j] * (t->xmax[j] - x[j]);
    mx = fmax(mx, fmax(fabs(rex), fmax(fabs(rexsi), fabs(reeta))));
    xs += rex * rex + rexsi * rexsi + reeta * reeta;
  }
  double la = 0.0;
  for (int i = 0; i < m; ++i) la += t->lambda[i] * t->a[i];
  const double rez = t->a0 - t->zeta - la;
  const double rezeta = t->zeta * t->z;
  for (int i = 0; i < m; ++i) {
    const double rey = t->c[i] + t->d[i] * t->y[i] - t->lambda[i] - t->mu[i];
    const double rel = fval[i] - t->a[i] * t->z - t->y[i] + t->s[i];
    const double remu = t->mu[i] * t->y[i];
    const double res = t->lambda[i] * t->s[i];
    mx = fmax(mx, fmax(fmax(fabs(rey), fabs(rel)), fmax(fabs(remu), fabs(res))));
    small += rey * rey + rel * rel + remu * remu + res * res;
  }
  mx = fmax(mx, fmax(fabs(rez), fabs(rezeta)));
  small += rez * rez + rezeta * rezeta;
  t->residumax = mx;
  t->residunorm = sqrt(small + xs);
}
