/*
 * pnpn.c -- FP64 CPU restatement of one Pn/Pn time step of the Brinkman
 * channel (the caller of ax_helm / gs / Jacobi-PCG), in the order of
 * sources/adjoint/adjoint_pnpn.f90:619-827:
 *   extrapolate velocity (sumab, :665-666); explicit forcing and advection
 *   (source_term%compute, opcolv B, adv, :669-682); makeabf / makebdf
 *   (:688-696); lag update (:698-700); pressure residual (:716-723), gs and
 *   bc (:725-726); pressure solve (:734-736), p += dp (:743); velocity
 *   residual (:747-753), gs and bc (:755-773); velocity solves (:784-786),
 *   u += du (:797-801).
 *
 * TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * The residual operators live in Neko (pnpn_prs_res / pnpn_vel_res, not
 * vendored); they are restated from the velocity-pressure splitting of
 * Tomboulides et al. (1997) with Neko's operator names:
 *   pressure : A_p dp = cdtp( Binv gs( F - mu B curl curl u_e - chi B u_e ) )
 *                       - A_p p - (rho bd0/dt) int_Gamma_D q u.n
 *   velocity : H du = F - opgrad(p) - H u,  H = mu A + (rho bd0/dt + chi) B
 * where F = EXT(B f) + (B rho/dt)(bd1 u^n + bd2 u^{n-1}) and
 * f = -(u.grad)u.  The Brinkman term is semi-implicit: chi is part of the
 * velocity Helmholtz operator (the explicit EXT2 form of the reference,
 * source_term%compute then makeabf, is unstable for chi dt > 4/3 and the
 * workloads have chi dt = 10).  Curl and the pressure forcing are
 * mass-averaged across elements (gs(B .) / gs(B)) as Neko's curl and
 * pnpn_prs_res do.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* ---- element metric and derivatives -------------------------------------*/
typedef struct {
  int lx;
  const double *D, *x, *y, *z, *w; /* w: GLL weights */
} geo_t;

/* r/s/t derivatives of f on element e (f points at the element block). */
static void elem_deriv(int lx, const double *D, const double *f, double *fr,
                       double *fs, double *ft) {
  for (int k = 0; k < lx; ++k)
    for (int j = 0; j < lx; ++j)
      for (int i = 0; i < lx; ++i) {
        double a = 0, b = 0, c = 0;
        for (int m = 0; m < lx; ++m) {
          a += D[i * lx + m] * f[m + lx * (j + lx * k)];
          b += D[j * lx + m] * f[i + lx * (m + lx * k)];
          c += D[k * lx + m] * f[i + lx * (j + lx * m)];
        }
        const int p = i + lx * (j + lx * k);
        fr[p] = a;
        fs[p] = b;
        ft[p] = c;
      }
}

/* cofactor rows c_r, c_s, c_t (9 per point) and J of element e. */
static void elem_cof(const geo_t *g, int64_t off, double *cof, double *J) {
  const int lx = g->lx, np = lx * lx * lx;
  double *buf = (double *)malloc(sizeof(double) * np * 9);
  double *xr = buf, *xs = buf + np, *xt = buf + 2 * np, *yr = buf + 3 * np,
         *ys = buf + 4 * np, *yt = buf + 5 * np, *zr = buf + 6 * np,
         *zs = buf + 7 * np, *zt = buf + 8 * np;
  elem_deriv(lx, g->D, g->x + off, xr, xs, xt);
  elem_deriv(lx, g->D, g->y + off, yr, ys, yt);
  elem_deriv(lx, g->D, g->z + off, zr, zs, zt);
  for (int p = 0; p < np; ++p) {
    const double ar[3] = {xr[p], yr[p], zr[p]}, as[3] = {xs[p], ys[p], zs[p]},
                 at[3] = {xt[p], yt[p], zt[p]};
    double *c = cof + 9 * p;
    c[0] = as[1] * at[2] - as[2] * at[1];
    c[1] = as[2] * at[0] - as[0] * at[2];
    c[2] = as[0] * at[1] - as[1] * at[0];
    c[3] = at[1] * ar[2] - at[2] * ar[1];
    c[4] = at[2] * ar[0] - at[0] * ar[2];
    c[5] = at[0] * ar[1] - at[1] * ar[0];
    c[6] = ar[1] * as[2] - ar[2] * as[1];
    c[7] = ar[2] * as[0] - ar[0] * as[2];
    c[8] = ar[0] * as[1] - ar[1] * as[0];
    J[p] = ar[0] * c[0] + ar[1] * c[1] + ar[2] * c[2];
  }
  free(buf);
}

/* The nine cofactor fields of every element, cof[m*n + l] with m = 3b + a
 * holding c_b[a] = J d(r_b)/d(x_a) (Neko coef drdx, drdy, .., dtdz). */
void or_cofactors(int lx, int nelv, const double *D, const double *x,
                  const double *y, const double *z, double *cof9) {
  const int np = lx * lx * lx;
  const int64_t n = (int64_t)nelv * np;
  geo_t g = {lx, D, x, y, z, NULL};
#pragma omp parallel num_threads(or_get_threads())
  {
    double *cof = (double *)malloc(sizeof(double) * np * 10);
    double *J = cof + 9 * np;
#pragma omp for schedule(static)
    for (int e = 0; e < nelv; ++e) {
      const int64_t o = (int64_t)e * np;
      elem_cof(&g, o, cof, J);
      for (int p = 0; p < np; ++p)
        for (int m = 0; m < 9; ++m) cof9[m * n + o + p] = cof[9 * p + m];
    }
    free(cof);
  }
}

/* -(u.grad) u_a  times B, per component: the convective term of the
 * explicit forcing (Neko adv_no_dealias, conv1 then -B*, per element). */
void or_advection(int lx, int nelv, const double *D, const double *x,
                  const double *y, const double *z, const double *B,
                  const double *u, const double *v, const double *w,
                  double *fx, double *fy, double *fz) {
  const int np = lx * lx * lx;
  geo_t g = {lx, D, x, y, z, NULL};
#pragma omp parallel num_threads(or_get_threads())
  {
    double *cof = (double *)malloc(sizeof(double) * np * 13);
    double *J = cof + 9 * np, *fr = J + np, *fs = fr + np, *ft = fs + np;
#pragma omp for schedule(static)
    for (int e = 0; e < nelv; ++e) {
      const int64_t o = (int64_t)e * np;
      elem_cof(&g, o, cof, J);
      const double *comp[3] = {u + o, v + o, w + o};
      double *out[3] = {fx + o, fy + o, fz + o};
      for (int a = 0; a < 3; ++a) {
        elem_deriv(lx, D, comp[a], fr, fs, ft);
        for (int p = 0; p < np; ++p) {
          const double *c = cof + 9 * p;
          const double gx = (c[0] * fr[p] + c[3] * fs[p] + c[6] * ft[p]) / J[p];
          const double gy = (c[1] * fr[p] + c[4] * fs[p] + c[7] * ft[p]) / J[p];
          const double gz = (c[2] * fr[p] + c[5] * fs[p] + c[8] * ft[p]) / J[p];
          const double conv = comp[0][p] * gx + comp[1][p] * gy + comp[2][p] * gz;
          out[a][p] = -B[o + p] * conv;
        }
      }
    }
    free(cof);
  }
}

/* Element-local curl (no averaging). */
void or_curl_local(int lx, int nelv, const double *D, const double *x,
                   const double *y, const double *z, const double *u,
                   const double *v, const double *w, double *cx, double *cy,
                   double *cz) {
  const int np = lx * lx * lx;
  geo_t g = {lx, D, x, y, z, NULL};
#pragma omp parallel num_threads(or_get_threads())
  {
    double *cof = (double *)malloc(sizeof(double) * np * 19);
    double *J = cof + 9 * np, *d = J + np; /* 9 derivative arrays */
#pragma omp for schedule(static)
    for (int e = 0; e < nelv; ++e) {
      const int64_t o = (int64_t)e * np;
      elem_cof(&g, o, cof, J);
      const double *comp[3] = {u + o, v + o, w + o};
      double gr[3][3]; /* not used as storage; see loop */
      (void)gr;
      for (int a = 0; a < 3; ++a)
        elem_deriv(lx, D, comp[a], d + (3 * a) * np, d + (3 * a + 1) * np,
                   d + (3 * a + 2) * np);
      for (int p = 0; p < np; ++p) {
        const double *c = cof + 9 * p;
        double G[3][3]; /* G[a][b] = d comp_a / d x_b */
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b)
            G[a][b] = (c[b] * d[(3 * a) * np + p] + c[3 + b] * d[(3 * a + 1) * np + p] +
                       c[6 + b] * d[(3 * a + 2) * np + p]) / J[p];
        cx[o + p] = G[2][1] - G[1][2];
        cy[o + p] = G[0][2] - G[2][0];
        cz[o + p] = G[1][0] - G[0][1];
      }
    }
    free(cof);
  }
}

/* opgrad: B grad p, per component (w3 sum_b c_b[a] p_b). */
void or_opgrad(int lx, int nelv, const double *D, const double *wq,
               const double *x, const double *y, const double *z,
               const double *p, double *ox, double *oy, double *oz) {
  const int np = lx * lx * lx;
  geo_t g = {lx, D, x, y, z, wq};
#pragma omp parallel num_threads(or_get_threads())
  {
    double *cof = (double *)malloc(sizeof(double) * np * 13);
    double *J = cof + 9 * np, *fr = J + np, *fs = fr + np, *ft = fs + np;
#pragma omp for schedule(static)
    for (int e = 0; e < nelv; ++e) {
      const int64_t o = (int64_t)e * np;
      elem_cof(&g, o, cof, J);
      elem_deriv(lx, D, p + o, fr, fs, ft);
      for (int k = 0; k < lx; ++k)
        for (int j = 0; j < lx; ++j)
          for (int i = 0; i < lx; ++i) {
            const int q = i + lx * (j + lx * k);
            const double w3 = wq[i] * wq[j] * wq[k];
            const double *c = cof + 9 * q;
            ox[o + q] = w3 * (c[0] * fr[q] + c[3] * fs[q] + c[6] * ft[q]);
            oy[o + q] = w3 * (c[1] * fr[q] + c[4] * fs[q] + c[7] * ft[q]);
            oz[o + q] = w3 * (c[2] * fr[q] + c[5] * fs[q] + c[8] * ft[q]);
          }
    }
    free(cof);
  }
}

/* cdtp summed over components: out = sum_b D_b^T ( w3 sum_a c_b[a] t_a ),
 * the transpose of opgrad (weak divergence). */
void or_cdtp(int lx, int nelv, const double *D, const double *wq,
             const double *x, const double *y, const double *z,
             const double *tx, const double *ty, const double *tz,
             double *out) {
  const int np = lx * lx * lx;
  geo_t g = {lx, D, x, y, z, wq};
#pragma omp parallel num_threads(or_get_threads())
  {
    double *cof = (double *)malloc(sizeof(double) * np * 13);
    double *J = cof + 9 * np, *gr = J + np, *gs = gr + np, *gt = gs + np;
#pragma omp for schedule(static)
    for (int e = 0; e < nelv; ++e) {
      const int64_t o = (int64_t)e * np;
      elem_cof(&g, o, cof, J);
      for (int k = 0; k < lx; ++k)
        for (int j = 0; j < lx; ++j)
          for (int i = 0; i < lx; ++i) {
            const int q = i + lx * (j + lx * k);
            const double w3 = wq[i] * wq[j] * wq[k];
            const double *c = cof + 9 * q;
            const double t0 = tx[o + q], t1 = ty[o + q], t2 = tz[o + q];
            gr[q] = w3 * (c[0] * t0 + c[1] * t1 + c[2] * t2);
            gs[q] = w3 * (c[3] * t0 + c[4] * t1 + c[5] * t2);
            gt[q] = w3 * (c[6] * t0 + c[7] * t1 + c[8] * t2);
          }
      for (int k = 0; k < lx; ++k)
        for (int j = 0; j < lx; ++j)
          for (int i = 0; i < lx; ++i) {
            double s = 0;
            for (int m = 0; m < lx; ++m) {
              s += D[m * lx + i] * gr[m + lx * (j + lx * k)];
              s += D[m * lx + j] * gs[i + lx * (m + lx * k)];
              s += D[m * lx + k] * gt[i + lx * (j + lx * m)];
            }
            out[o + i + lx * (j + lx * k)] = s;
          }
    }
    free(cof);
  }
}

/* Boundary flux: out(pt) = sum over the box faces in `faces` containing the
 * point of  w w (+-c_face . u(pt)),  i.e. int_face q u.n for the GLL
 * basis function q of the point (outward normal; c_face the cofactor row of
 * the face's reference direction).  Box faces: bit0 x=0 (r=-1 of ex=0), bit1
 * x=L (r=+1 of ex=nx-1), bit2/3 y, bit4/5 z. */
void or_surface_flux(int nx, int ny, int nz, int lx, const double *D,
                     const double *wq, const double *x, const double *y,
                     const double *z, const double *u, const double *v,
                     const double *w, int faces, double *out) {
  const int np = lx * lx * lx, P = lx - 1;
  const int nelv = nx * ny * nz;
  geo_t g = {lx, D, x, y, z, wq};
  double *cof = (double *)malloc(sizeof(double) * np * 10);
  double *J = cof + 9 * np;
  memset(out, 0, sizeof(double) * (size_t)nelv * np);
  for (int e = 0; e < nelv; ++e) {
    const int ex = e % nx, ey = (e / nx) % ny, ez = e / (nx * ny);
    const int on[6] = {(faces & 1) && ex == 0, (faces & 2) && ex == nx - 1,
                       (faces & 4) && ey == 0, (faces & 8) && ey == ny - 1,
                       (faces & 16) && ez == 0, (faces & 32) && ez == nz - 1};
    if (!(on[0] | on[1] | on[2] | on[3] | on[4] | on[5])) continue;
    const int64_t o = (int64_t)e * np;
    elem_cof(&g, o, cof, J);
    for (int f = 0; f < 6; ++f) {
      if (!on[f]) continue;
      const int dir = f / 2;              /* 0: r, 1: s, 2: t */
      const int hi = f % 2;               /* +1 face */
      const double sgn = hi ? 1.0 : -1.0;
      for (int b = 0; b < lx; ++b)
        for (int a = 0; a < lx; ++a) {
          int i, j, k;
          double wt;
          if (dir == 0) { i = hi ? P : 0; j = a; k = b; wt = wq[j] * wq[k]; }
          else if (dir == 1) { j = hi ? P : 0; i = a; k = b; wt = wq[i] * wq[k]; }
          else { k = hi ? P : 0; i = a; j = b; wt = wq[i] * wq[j]; }
          const int q = i + lx * (j + lx * k);
          const double *c = cof + 9 * q + 3 * dir;
          const double un = c[0] * u[o + q] + c[1] * v[o + q] + c[2] * w[o + q];
          out[o + q] += sgn * wt * un;
        }
    }
  }
  free(cof);
}

/* Adjoint advection without dealiasing, the device branch of
 * adjoint_advection_no_dealias (sources/adjoint/adv_adjoint_no_dealias.f90:
 * 119-255 with adjoint_weak_no_dealias_device :269-303), for the adjoint
 * velocity (ua, va, wa) about the base flow (ub, vb, wb):
 *   f_c -= sum_d ua_d (B d ub_d / dx_c)                  (opgrad, :208-219)
 *   f_c -= sum_d cdtp_d(ua_c ub_d)                        (:293-302)
 * The reference's CPU branch (:331-345) adds w2 and w3 instead of subtracting
 * them; the device branch is the discrete adjoint of the linearised
 * advection (tests/test_oracle_pnpn.py checks the duality), so it is the one
 * restated.  work: 12 n doubles. */
void or_adjoint_advection(int lx, int nelv, const double *D, const double *wq,
                          const double *x, const double *y, const double *z,
                          const double *ua, const double *va, const double *wa,
                          const double *ub, const double *vb, const double *wb,
                          double *fx, double *fy, double *fz, double *work) {
  const int64_t n = (int64_t)nelv * lx * lx * lx;
  double *g = work, *pr = work + 9 * n;
  or_opgrad(lx, nelv, D, wq, x, y, z, ub, g, g + n, g + 2 * n);
  or_opgrad(lx, nelv, D, wq, x, y, z, vb, g + 3 * n, g + 4 * n, g + 5 * n);
  or_opgrad(lx, nelv, D, wq, x, y, z, wb, g + 6 * n, g + 7 * n, g + 8 * n);
  for (int64_t i = 0; i < n; ++i) {
    fx[i] -= ua[i] * g[i] + va[i] * g[3 * n + i] + wa[i] * g[6 * n + i];
    fy[i] -= ua[i] * g[n + i] + va[i] * g[4 * n + i] + wa[i] * g[7 * n + i];
    fz[i] -= ua[i] * g[2 * n + i] + va[i] * g[5 * n + i] + wa[i] * g[8 * n + i];
  }
  const double *uc[3] = {ua, va, wa};
  double *fc[3] = {fx, fy, fz};
  for (int c = 0; c < 3; ++c) {
    for (int64_t i = 0; i < n; ++i) {
      pr[i] = uc[c][i] * ub[i];
      pr[n + i] = uc[c][i] * vb[i];
      pr[2 * n + i] = uc[c][i] * wb[i];
    }
    or_cdtp(lx, nelv, D, wq, x, y, z, pr, pr + n, pr + 2 * n, g);
    for (int64_t i = 0; i < n; ++i) fc[c][i] -= g[i];
  }
}

/* ---- the step ------------------------------------------------------------*/
typedef struct {
  int nx, ny, nz, lx, nelv;
  int64_t n;
  const double *D, *wq, *x, *y, *z, *B, *Binv, *mult, *chi;
  const double *G11, *G12, *G13, *G22, *G23, *G33;
  int gs_nb;
  const int *gs_m;
  const int64_t *gs_cnt;
  const int32_t *gs_idx;
  double volume;
  int64_t nmask_v, nmask_p;
  const int32_t *mask_v, *mask_p;
  int faces_v; /* Dirichlet velocity faces (surface flux) */
  double Re, dt, rho, abstol_p, abstol_v;
  int max_iter;
  /* state */
  double *u, *v, *w, *p, *ulag, *vlag, *wlag, *fxlag, *fylag, *fzlag;
  int tstep;
  /* per-step monitor: iters[0] pressure, iters[1..3] velocity */
  int iters[4];
  double res[4];
  /* adjoint mode: the state is the adjoint velocity and pressure about the
   * frozen base flow ub, vb, wb (adjoint_pnpn_step, adjoint_pnpn.f90:619-827) */
  int adjoint;
  const double *ub, *vb, *wb;
  double K, obj_scale;
  /* dealiased advection (case.numerics.dealias): lxd GL points, the GLL->GL
   * interpolation Jm, GL derivative matrix Dg, GL weights wg, GL metric cof
   * (dealias.c); dealias = 0 keeps the collocated (no-dealias) forms */
  int dealias, lxd;
  const double *Jm, *Dg, *wg, *cof;
  /* planted mistakes (OR_PLANT_* bits; 0 = the restatement as written): the
   * oracle tests switch one on to show that the exact-solution and
   * convergence checks of the step fail for it */
  int plant;
  /* Dirichlet velocity values at the new time level (full fields, read on
   * mask_v only; NULL = keep the values u, v, w hold), set after the lag
   * update as Neko's bc_apply_vel(t, tstep) does before the residuals */
  const double *ubc, *vbc, *wbc;
  /* 1: the reference's explicit Brinkman term (source_term%compute, then
   * opcolv B and EXT via makeabf, adjoint_pnpn.f90:669-691): f -= chi B u
   * after the advection, and chi is in neither the velocity Helmholtz
   * operator nor the pressure forcing; 0: semi-implicit */
  int brinkman_explicit;
  /* case.numerics.time_order: 2 (BDF2/EXT2 after a BDF1/EXT1 first step;
   * also when 0) or 3 (then BDF2/EXT2 on step 2 and BDF3/EXT3 from step 3
   * on, Neko's ext_bdf ramp); order 3 needs the second lags below */
  int time_order;
  double *ulag2, *vlag2, *wlag2, *fxlag2, *fylag2, *fzlag2;
} or_pnpn_t;

enum {
  OR_PLANT_MU2 = 1,            /* viscosity doubled everywhere */
  OR_PLANT_CURLCURL_SIGN = 2,  /* + mu B curl curl u_e in the pressure forcing */
  OR_PLANT_SWAP_BDF_LAGS = 4,  /* bd1 applied to u^{n-1}, bd2 to u^n */
  OR_PLANT_SURF_DT = 8,        /* surface term with rho/dt, not rho bd0/dt */
  OR_PLANT_ADV_SIGN = 16,      /* +B (u.grad) u as the explicit forcing */
  OR_PLANT_CHI_PRS_OFF = 32,   /* chi B u_e left out of the pressure forcing */
  OR_PLANT_EXT_SWAP = 64,      /* EXT coefficients in swapped order */
  OR_PLANT_CHI_VEL_OFF = 128,  /* chi left out of the velocity Helmholtz h2 */
  OR_PLANT_BDF1_ONLY = 256     /* BDF1/EXT1 on every step */
};

static void avg_field(const or_pnpn_t *s, double *a) {
  for (int64_t i = 0; i < s->n; ++i) a[i] = a[i] * s->B[i];
  or_gs_op(a, s->gs_nb, s->gs_m, s->gs_cnt, s->gs_idx, OR_GS_ADD);
  for (int64_t i = 0; i < s->n; ++i) a[i] = a[i] * s->Binv[i];
}

/* BDF/EXT coefficients for a constant step (Neko ext_bdf_t, not vendored):
 * order k = min(tstep, time_order); BDF_k: bd[0] u^{n+1} = sum_{j>=1}
 * bd[j] u^{n+1-j} + dt (...), EXT_k: u^{n+1} ~ sum_j ext[j] u^{n-j}.
 *   k = 1: bd = {1, 1},           ext = {1}
 *   k = 2: bd = {3/2, 2, -1/2},   ext = {2, -1}
 *   k = 3: bd = {11/6, 3, -3/2, 1/3}, ext = {3, -3, 1}
 * (the backward-difference and extrapolation formulas of the Lagrange
 * polynomial through the last k + 1, resp. k, levels). */
void or_pnpn_coeffs3(int tstep, int order, double *bd, double *ext) {
  const int k = tstep < order ? tstep : order;
  bd[0] = bd[1] = bd[2] = bd[3] = 0.0;
  ext[0] = ext[1] = ext[2] = 0.0;
  if (k <= 1) {
    bd[0] = 1.0; bd[1] = 1.0;
    ext[0] = 1.0;
  } else if (k == 2) {
    bd[0] = 1.5; bd[1] = 2.0; bd[2] = -0.5;
    ext[0] = 2.0; ext[1] = -1.0;
  } else {
    bd[0] = 11.0 / 6.0; bd[1] = 3.0; bd[2] = -1.5; bd[3] = 1.0 / 3.0;
    ext[0] = 3.0; ext[1] = -3.0; ext[2] = 1.0;
  }
}

void or_pnpn_coeffs(int tstep, double *bd, double *ext) {
  double b[4], e[3];
  or_pnpn_coeffs3(tstep, 2, b, e);
  bd[0] = b[0]; bd[1] = b[1]; bd[2] = b[2];
  ext[0] = e[0]; ext[1] = e[1];
}

void or_pnpn_step(or_pnpn_t *s) {
  const int64_t n = s->n;
  const int lx = s->lx, nelv = s->nelv;
  const int pl = s->plant;
  const double mu = ((pl & OR_PLANT_MU2) ? 2.0 : 1.0) / s->Re, rho = s->rho,
               dt = s->dt;
  s->tstep += 1;
  const int order = s->time_order == 3 ? 3 : 2;
  double bd[4], ext[3];
  or_pnpn_coeffs3((pl & OR_PLANT_BDF1_ONLY) ? 1 : s->tstep, order, bd, ext);
  const int o3 = order == 3;
  if (pl & OR_PLANT_EXT_SWAP) {
    const double e0 = ext[0];
    ext[0] = ext[1];
    ext[1] = e0;
  }
  const double bdu = (pl & OR_PLANT_SWAP_BDF_LAGS) ? bd[2] : bd[1];
  const double bdl = (pl & OR_PLANT_SWAP_BDF_LAGS) ? bd[1] : bd[2];
  double *buf = (double *)calloc((size_t)n * 20, sizeof(double));
  double *ue = buf, *ve = buf + n, *we = buf + 2 * n;
  double *fx = buf + 3 * n, *fy = buf + 4 * n, *fz = buf + 5 * n;
  double *Fx = buf + 6 * n, *Fy = buf + 7 * n, *Fz = buf + 8 * n;
  double *t1 = buf + 9 * n, *t2 = buf + 10 * n, *t3 = buf + 11 * n;
  double *c1 = buf + 12 * n, *c2 = buf + 13 * n, *c3 = buf + 14 * n;
  double *pres = buf + 15 * n, *h1 = buf + 16 * n, *h2 = buf + 17 * n,
         *dinv = buf + 18 * n, *dx = buf + 19 * n;
  or_sem_t sem = {lx, nelv, n, s->D, s->G11, s->G12, s->G13, s->G22, s->G23,
                  s->G33, s->B, s->mult, s->gs_nb, s->gs_m, s->gs_cnt,
                  s->gs_idx, 0, NULL, s->volume};
  /* (1) extrapolated velocity */
  for (int64_t i = 0; i < n; ++i) {
    ue[i] = ext[0] * s->u[i] + ext[1] * s->ulag[i];
    ve[i] = ext[0] * s->v[i] + ext[1] * s->vlag[i];
    we[i] = ext[0] * s->w[i] + ext[1] * s->wlag[i];
    if (o3) {
      ue[i] += ext[2] * s->ulag2[i];
      ve[i] += ext[2] * s->vlag2[i];
      we[i] += ext[2] * s->wlag2[i];
    }
  }
  if (!s->adjoint) {
    /* (2) explicit forcing B f = -B (u.grad) u; dealiased: f = 0 - J^T[..]
     * (calloc'd), Neko adv_dealias_t%compute */
    if (s->dealias)
      or_dealias_advection(lx, s->lxd, nelv, s->Jm, s->Dg, s->wg, s->cof, s->u,
                           s->v, s->w, fx, fy, fz);
    else
      or_advection(lx, nelv, s->D, s->x, s->y, s->z, s->B, s->u, s->v, s->w,
                   fx, fy, fz);
    if (pl & OR_PLANT_ADV_SIGN)
      for (int64_t i = 0; i < n; ++i) {
        fx[i] = -fx[i];
        fy[i] = -fy[i];
        fz[i] = -fz[i];
      }
  } else {
    /* (2) adjoint forcing (adjoint_pnpn.f90:669-681):
     *   source terms f = obj_scale curl(curl ub) (Neko curl: mass-averaged,
     *   adjoint_minimum_dissipation_source_term.f90:229-243) + K chi ub
     *   (adjoint_lube_source_term.f90:189-203); f *= B (opcolv); then the
     *   adjoint advection about ub. */
    or_curl_local(lx, nelv, s->D, s->x, s->y, s->z, s->ub, s->vb, s->wb, t1, t2,
                  t3);
    avg_field(s, t1);
    avg_field(s, t2);
    avg_field(s, t3);
    or_curl_local(lx, nelv, s->D, s->x, s->y, s->z, t1, t2, t3, c1, c2, c3);
    avg_field(s, c1);
    avg_field(s, c2);
    avg_field(s, c3);
    for (int64_t i = 0; i < n; ++i) {
      const double kc = s->K * s->chi[i];
      fx[i] = (s->obj_scale * c1[i] + kc * s->ub[i]) * s->B[i];
      fy[i] = (s->obj_scale * c2[i] + kc * s->vb[i]) * s->B[i];
      fz[i] = (s->obj_scale * c3[i] + kc * s->wb[i]) * s->B[i];
    }
    if (s->dealias) {
      or_dealias_adjoint(lx, s->lxd, nelv, s->Jm, s->Dg, s->wg, s->cof, s->u,
                         s->v, s->w, s->ub, s->vb, s->wb, fx, fy, fz);
    } else {
      double *work = (double *)malloc(sizeof(double) * 12 * (size_t)n);
      or_adjoint_advection(lx, nelv, s->D, s->wq, s->x, s->y, s->z, s->u, s->v,
                           s->w, s->ub, s->vb, s->wb, fx, fy, fz, work);
      free(work);
    }
  }
  if (s->brinkman_explicit)
    for (int64_t i = 0; i < n; ++i) {
      const double c = s->chi[i] * s->B[i];
      fx[i] = fx[i] - s->u[i] * c;
      fy[i] = fy[i] - s->v[i] * c;
      fz[i] = fz[i] - s->w[i] * c;
    }
  /* (3) makeabf and (4) makebdf; lag update */
  for (int64_t i = 0; i < n; ++i) {
    const double br = s->B[i] * rho / dt;
    Fx[i] = ext[0] * fx[i] + ext[1] * s->fxlag[i] + br * (bdu * s->u[i] + bdl * s->ulag[i]);
    Fy[i] = ext[0] * fy[i] + ext[1] * s->fylag[i] + br * (bdu * s->v[i] + bdl * s->vlag[i]);
    Fz[i] = ext[0] * fz[i] + ext[1] * s->fzlag[i] + br * (bdu * s->w[i] + bdl * s->wlag[i]);
    if (o3) {
      Fx[i] += ext[2] * s->fxlag2[i] + br * bd[3] * s->ulag2[i];
      Fy[i] += ext[2] * s->fylag2[i] + br * bd[3] * s->vlag2[i];
      Fz[i] += ext[2] * s->fzlag2[i] + br * bd[3] * s->wlag2[i];
      s->fxlag2[i] = s->fxlag[i];
      s->fylag2[i] = s->fylag[i];
      s->fzlag2[i] = s->fzlag[i];
      s->ulag2[i] = s->ulag[i];
      s->vlag2[i] = s->vlag[i];
      s->wlag2[i] = s->wlag[i];
    }
    s->fxlag[i] = fx[i];
    s->fylag[i] = fy[i];
    s->fzlag[i] = fz[i];
    s->ulag[i] = s->u[i];
    s->vlag[i] = s->v[i];
    s->wlag[i] = s->w[i];
  }
  if (s->ubc)
    for (int64_t q = 0; q < s->nmask_v; ++q) {
      const int32_t i = s->mask_v[q];
      s->u[i] = s->ubc[i];
      s->v[i] = s->vbc[i];
      s->w[i] = s->wbc[i];
    }
  /* (5) pressure residual */
  or_curl_local(lx, nelv, s->D, s->x, s->y, s->z, ue, ve, we, t1, t2, t3);
  avg_field(s, t1);
  avg_field(s, t2);
  avg_field(s, t3);
  or_curl_local(lx, nelv, s->D, s->x, s->y, s->z, t1, t2, t3, c1, c2, c3);
  const double ccs = (pl & OR_PLANT_CURLCURL_SIGN) ? -1.0 : 1.0;
  const double chp = ((pl & OR_PLANT_CHI_PRS_OFF) || s->brinkman_explicit) ? 0.0 : 1.0;
  for (int64_t i = 0; i < n; ++i) {
    const double mb = ccs * mu * s->B[i], cb = chp * s->chi[i] * s->B[i];
    t1[i] = Fx[i] / rho - mb * c1[i] - cb * ue[i];
    t2[i] = Fy[i] / rho - mb * c2[i] - cb * ve[i];
    t3[i] = Fz[i] / rho - mb * c3[i] - cb * we[i];
  }
  or_gs_op(t1, s->gs_nb, s->gs_m, s->gs_cnt, s->gs_idx, OR_GS_ADD);
  or_gs_op(t2, s->gs_nb, s->gs_m, s->gs_cnt, s->gs_idx, OR_GS_ADD);
  or_gs_op(t3, s->gs_nb, s->gs_m, s->gs_cnt, s->gs_idx, OR_GS_ADD);
  for (int64_t i = 0; i < n; ++i) {
    t1[i] *= s->Binv[i];
    t2[i] *= s->Binv[i];
    t3[i] *= s->Binv[i];
  }
  or_cdtp(lx, nelv, s->D, s->wq, s->x, s->y, s->z, t1, t2, t3, pres);
  for (int64_t i = 0; i < n; ++i) {
    h1[i] = 1.0 / rho;
    h2[i] = 0.0;
  }
  or_ax_helm(lx, nelv, c1, s->p, s->D, h1, h2, s->G11, s->G12, s->G13, s->G22,
             s->G23, s->G33, s->B, 0);
  or_surface_flux(s->nx, s->ny, s->nz, lx, s->D, s->wq, s->x, s->y, s->z,
                  s->u, s->v, s->w, s->faces_v, c2);
  const double dtbd = rho * ((pl & OR_PLANT_SURF_DT) ? 1.0 : bd[0]) / dt;
  for (int64_t i = 0; i < n; ++i) pres[i] = pres[i] - c1[i] - dtbd * c2[i];
  or_gs_op(pres, s->gs_nb, s->gs_m, s->gs_cnt, s->gs_idx, OR_GS_ADD);
  for (int64_t q = 0; q < s->nmask_p; ++q) pres[s->mask_p[q]] = 0.0;
  /* (6) pressure solve, p += dp */
  sem.nmask = s->nmask_p;
  sem.mask = s->mask_p;
  or_jacobi_update(&sem, h1, h2, 0, dinv);
  or_cg_solve(&sem, h1, h2, 0, dinv, dx, pres, s->abstol_p, s->max_iter,
              &s->iters[0], &s->res[0], &s->res[0]);
  for (int64_t i = 0; i < n; ++i) s->p[i] += dx[i];
  /* (7) velocity residual */
  for (int64_t i = 0; i < n; ++i) {
    h1[i] = mu;
    h2[i] = rho * bd[0] / dt +
            (((pl & OR_PLANT_CHI_VEL_OFF) || s->brinkman_explicit) ? 0.0 : s->chi[i]);
  }
  or_opgrad(lx, nelv, s->D, s->wq, s->x, s->y, s->z, s->p, t1, t2, t3);
  double *uu[3] = {s->u, s->v, s->w}, *FF[3] = {Fx, Fy, Fz}, *tt[3] = {t1, t2, t3};
  double *rr[3] = {c1, c2, c3};
  for (int a = 0; a < 3; ++a) {
    or_ax_helm(lx, nelv, rr[a], uu[a], s->D, h1, h2, s->G11, s->G12, s->G13,
               s->G22, s->G23, s->G33, s->B, 1);
    for (int64_t i = 0; i < n; ++i) rr[a][i] = FF[a][i] - tt[a][i] - rr[a][i];
    or_gs_op(rr[a], s->gs_nb, s->gs_m, s->gs_cnt, s->gs_idx, OR_GS_ADD);
    for (int64_t q = 0; q < s->nmask_v; ++q) rr[a][s->mask_v[q]] = 0.0;
  }
  /* (8) velocity solves, u += du */
  sem.nmask = s->nmask_v;
  sem.mask = s->mask_v;
  or_jacobi_update(&sem, h1, h2, 1, dinv);
  for (int a = 0; a < 3; ++a) {
    double r0;
    or_cg_solve(&sem, h1, h2, 1, dinv, dx, rr[a], s->abstol_v, s->max_iter,
                &s->iters[1 + a], &r0, &s->res[1 + a]);
    for (int64_t i = 0; i < n; ++i) uu[a][i] += dx[i];
  }
  free(buf);
}
