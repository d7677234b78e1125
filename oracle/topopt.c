/*
 * topopt.c -- FP64 CPU restatement of neko-top's topology-optimisation
 * kernels on the hot path: Brinkman source term, the legacy permeability
 * force, RAMP and linear design maps with their chain rules, the PDE
 * (Helmholtz) filter, the minimum-dissipation objective with its "lube"
 * term and sensitivity, the adjoint lube forcing, and the volume
 * constraint.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle).  Operation order follows the cited
 * reference lines so FP64 results can be compared at 1e-12.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* f -= chi * u for each component.
 * sources/source_terms/simple_brinkman_source_term.f90:139-153
 * (field_subcol3(fu, u, chi): fu = fu - u*chi). */
void or_brinkman(int64_t n, double *fx, double *fy, double *fz,
                 const double *u, const double *v, const double *w,
                 const double *chi) {
  for (int64_t i = 0; i < n; ++i) {
    fx[i] = fx[i] - u[i] * chi[i];
    fy[i] = fy[i] - v[i] * chi[i];
    fz[i] = fz[i] - w[i] * chi[i];
  }
}

/* Legacy permeability: resistance = -(k1 + (k0-k1)(q+1) rho/(q+rho)),
 * built by the reference as cfill(q), add2(rho), invcol1, col2(rho),
 * cmult((k0-k1)(q+1)), cadd(k1), cmult(-1).
 * sources/topology_optimization/design.f90:316-348. */
void or_update_permeability(int64_t n, double *res, const double *rho,
                            double perm_0, double perm_1, double penalty) {
  const double constant = (perm_0 - perm_1) * (penalty + 1.0);
  for (int64_t i = 0; i < n; ++i) {
    double r = penalty;
    r = r + rho[i];
    r = 1.0 / r;
    r = r * rho[i];
    r = constant * r;
    r = r + perm_1;
    res[i] = -1.0 * r;
  }
}

/* f = resistance * u on the 1-based mask (col3_mask).
 * sources/topology_optimization/design.f90:266-294,
 * sources/neko_ext/math/math_ext.f90:61-71. */
void or_col3_mask(double *a, const double *b, const double *c,
                  const int *mask, int mask_size) {
  for (int i = 0; i < mask_size; ++i) {
    const int m = mask[i] - 1;
    a[m] = b[m] * c[m];
  }
}

/* RAMP forward.  convex_up:   f_min + (f_max-f_min) x (1+q)/(x+q)
 *                convex_down: f_min + (f_max-f_min) x / (1 + q (1-x))
 * sources/mapping_functions/RAMP_mapping.f90:182-196, :227-241. */
void or_ramp_fwd(int64_t n, double *out, const double *in, double f_min,
                 double f_max, double q, int convex_up) {
  for (int64_t i = 0; i < n; ++i) {
    const double x = in[i];
    if (convex_up)
      out[i] = f_min + (f_max - f_min) * x * (1.0 + q) / (x + q);
    else
      out[i] = f_min + (f_max - f_min) * x / (1.0 + q * (1.0 - x));
  }
}

/* RAMP chain rule.  convex_up:   (f_max-f_min)(q+1)/(x+q)^2 dF/dout
 *                   convex_down: (f_max-f_min)(q+1)/(1-q(x-1))^2 dF/dout
 * sources/mapping_functions/RAMP_mapping.f90:203-222, :248-267. */
void or_ramp_bwd(int64_t n, double *df_din, const double *df_dout,
                 const double *in, double f_min, double f_max, double q,
                 int convex_up) {
  for (int64_t i = 0; i < n; ++i) {
    const double x = in[i];
    double den;
    if (convex_up)
      den = (x + q) * (x + q);
    else
      den = (1.0 - q * (x - 1.0)) * (1.0 - q * (x - 1.0));
    df_din[i] = (f_max - f_min) * (q + 1.0) / den * df_dout[i];
  }
}

/* Linear map x_out = f_min + (f_max-f_min) x_in and its chain rule.
 * sources/mapping_functions/linear_mapping.f90:104-136. */
void or_linear_fwd(int64_t n, double *out, const double *in, double f_min,
                   double f_max) {
  for (int64_t i = 0; i < n; ++i) out[i] = (f_max - f_min) * in[i] + f_min;
}

void or_linear_bwd(int64_t n, double *df_din, const double *df_dout,
                   double f_min, double f_max) {
  for (int64_t i = 0; i < n; ++i) df_din[i] = (f_max - f_min) * df_dout[i];
}

/* Element-local gradient (Neko grad / opgrad without gs):
 * du/dx_a = (c_r[a] u_r + c_s[a] u_s + c_t[a] u_t) / J with the cofactors
 * c_r = x_s x x_t, c_s = x_t x x_r, c_t = x_r x x_s of the coordinates. */
void or_grad(int lx, int nelv, const double *D, const double *x,
             const double *y, const double *z, const double *u, double *ux,
             double *uy, double *uz) {
  const int np = lx * lx * lx;
#pragma omp parallel for num_threads(or_get_threads()) schedule(static)
  for (int e = 0; e < nelv; ++e) {
    const int64_t o = (int64_t)e * np;
    const double *f[4] = {x + o, y + o, z + o, u + o};
    for (int k = 0; k < lx; ++k)
      for (int j = 0; j < lx; ++j)
        for (int i = 0; i < lx; ++i) {
          double dr[4] = {0, 0, 0, 0}, ds[4] = {0, 0, 0, 0}, dt[4] = {0, 0, 0, 0};
          for (int a = 0; a < 4; ++a)
            for (int m = 0; m < lx; ++m) {
              dr[a] += D[i * lx + m] * f[a][m + lx * (j + lx * k)];
              ds[a] += D[j * lx + m] * f[a][i + lx * (m + lx * k)];
              dt[a] += D[k * lx + m] * f[a][i + lx * (j + lx * m)];
            }
          const double cr[3] = {ds[1] * dt[2] - ds[2] * dt[1],
                                ds[2] * dt[0] - ds[0] * dt[2],
                                ds[0] * dt[1] - ds[1] * dt[0]};
          const double cs[3] = {dt[1] * dr[2] - dt[2] * dr[1],
                                dt[2] * dr[0] - dt[0] * dr[2],
                                dt[0] * dr[1] - dt[1] * dr[0]};
          const double ct[3] = {dr[1] * ds[2] - dr[2] * ds[1],
                                dr[2] * ds[0] - dr[0] * ds[2],
                                dr[0] * ds[1] - dr[1] * ds[0]};
          const double J = dr[0] * cr[0] + dr[1] * cr[1] + dr[2] * cr[2];
          const int64_t g = o + i + lx * (j + lx * k);
          ux[g] = (cr[0] * dr[3] + cs[0] * ds[3] + ct[0] * dt[3]) / J;
          uy[g] = (cr[1] * dr[3] + cs[1] * ds[3] + ct[1] * dt[3]) / J;
          uz[g] = (cr[2] * dr[3] + cs[2] * ds[3] + ct[2] * dt[3]) / J;
        }
  }
}

/* Minimum-dissipation objective.
 * sources/objectives/minimum_dissipation_objective_function.f90:186-254:
 *   diss = glsc2(|grad u|^2 + |grad v|^2 + |grad w|^2, B)
 *   lube = glsc2(chi (u + v + w), B)       (sic: chi times the sum)
 *   F = (diss + 0.5 K lube) * obj_scale   (lube only if if_lube)
 * out[0] = F, out[1] = diss, out[2] = lube. */
void or_dissipation_objective(int lx, int nelv, const double *D,
                              const double *x, const double *y,
                              const double *z, const double *B,
                              const double *u, const double *v,
                              const double *w, const double *chi, double K,
                              double obj_scale, int if_lube, double *out) {
  const int64_t n = (int64_t)nelv * lx * lx * lx;
  double *g = (double *)malloc(sizeof(double) * n * 4);
  double *gx = g, *gy = g + n, *gz = g + 2 * n, *of = g + 3 * n;
  const double *comp[3] = {u, v, w};
  for (int c = 0; c < 3; ++c) {
    or_grad(lx, nelv, D, x, y, z, comp[c], gx, gy, gz);
    for (int64_t i = 0; i < n; ++i) {
      double s = (c == 0) ? gx[i] * gx[i] : of[i] + gx[i] * gx[i];
      s = s + gy[i] * gy[i];
      s = s + gz[i] * gz[i];
      of[i] = s;
    }
  }
  const double diss = or_glsc2(n, of, B);
  double lube = 0.0, F = diss;
  if (if_lube) {
    for (int64_t i = 0; i < n; ++i) {
      double s = u[i] * chi[i];
      s = s + v[i] * chi[i];
      s = s + w[i] * chi[i];
      of[i] = s;
    }
    lube = or_glsc2(n, of, B);
    F = diss + 0.5 * K * lube;
  }
  out[0] = F * obj_scale;
  out[1] = diss;
  out[2] = lube;
  free(g);
}

/* dF/dchi = -(u u_adj + v v_adj + w w_adj) [+ K obj_scale (u^2+v^2+w^2)].
 * sources/objectives/minimum_dissipation_objective_function.f90:260-301. */
void or_dissipation_sensitivity(int64_t n, double *s, const double *u,
                                const double *v, const double *w,
                                const double *ua, const double *va,
                                const double *wa, double K, double obj_scale,
                                int if_lube) {
  const double kk = K * obj_scale;
  for (int64_t i = 0; i < n; ++i) {
    double a = u[i] * ua[i];
    a = a + v[i] * va[i];
    a = a + w[i] * wa[i];
    a = -1.0 * a;
    if (if_lube) {
      double l = u[i] * u[i];
      l = l + v[i] * v[i];
      l = l + w[i] * w[i];
      a = a + kk * l;
    }
    s[i] = a;
  }
}

/* Adjoint lube forcing f_adj += (K chi) u (primal u), per component.
 * sources/source_terms/adjoint_lube_source_term.f90:173-206. */
void or_adjoint_lube(int64_t n, double *fx, double *fy, double *fz,
                     const double *u, const double *v, const double *w,
                     const double *chi, double K) {
  for (int64_t i = 0; i < n; ++i) {
    const double k = chi[i] * K;
    fx[i] = fx[i] + u[i] * k;
    fy[i] = fy[i] + v[i] * k;
    fz[i] = fz[i] + w[i] * k;
  }
}

/* Volume constraint on the unfiltered design.
 * sources/objectives/volume_constraint.f90:177-218:
 *   V = glsc2(rho, B) / volume_domain;  value = is_max ? V - v_max
 *                                                      : -V + v_max
 * out[0] = value, out[1] = V. */
void or_volume_constraint(int64_t n, const double *rho, const double *B,
                          double volume_domain, double v_max, int is_max,
                          double *out) {
  const double V = or_glsc2(n, rho, B) / volume_domain;
  out[1] = V;
  out[0] = is_max ? V - v_max : -V + v_max;
}

/* Its sensitivity: the constant field +-1/volume_domain.
 * sources/objectives/volume_constraint.f90:224-242. */
void or_volume_sensitivity(int64_t n, double *s, double volume_domain,
                           int is_max) {
  const double c = is_max ? 1.0 / volume_domain : -1.0 / volume_domain;
  for (int64_t i = 0; i < n; ++i) s[i] = 1.0 * c;
}

/* PDE filter -r^2 lap(y) + y = x with homogeneous Neumann boundaries:
 * h1 = r^2, h2 = 1, RHS = gs(B x) (forward; the adjoint applies bc then gs,
 * identical without Dirichlet points), solved by Jacobi-PCG from zero to
 * abstol.  sources/mapping_functions/PDE_filter_mapping.f90:212-363 (the
 * reference solves with GMRES + identity preconditioner; the solution of
 * the SPD system is the same to the tolerance). Returns the iterations. */
int or_filter_apply(const or_sem_t *s, double r, double abstol, int max_iter,
                    const double *in, double *out) {
  const int64_t n = s->n;
  double *buf = (double *)malloc(sizeof(double) * n * 4);
  double *h1 = buf, *h2 = buf + n, *rhs = buf + 2 * n, *dinv = buf + 3 * n;
  for (int64_t i = 0; i < n; ++i) {
    h1[i] = r * r;
    h2[i] = 1.0;
    rhs[i] = in[i] * s->B[i];
  }
  or_gs_op(rhs, s->gs_nb, s->gs_m, s->gs_cnt, s->gs_idx, OR_GS_ADD);
  or_sem_t t = *s;
  t.nmask = 0;
  or_jacobi_update(&t, h1, h2, 1, dinv);
  int iter;
  double r0, r1;
  or_cg_solve(&t, h1, h2, 1, dinv, out, rhs, abstol, max_iter, &iter, &r0,
              &r1);
  free(buf);
  return iter;
}
