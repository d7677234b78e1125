/*
 * oracle.h -- FP64 CPU restatement of the neko-top spectral-element hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; only the test-suite, the smoke check and the CPU-baseline
 * leg of the benchmark use it, and only as the checker.
 *
 * What it restates, and from where:
 *   - Brinkman source term, RAMP map, PDE filter, objective, volume
 *     constraint, MMA (gensub / subsolve / KKT): neko-top sources, cited per
 *     function in the .c files.
 *   - GLL space, box mesh, dofmap ids, geometric factors, ax_helm, gs_op,
 *     Jacobi and CG: these live in Neko (not vendored with neko-top).  They
 *     are restated from the textbook SEM formulation (Deville, Fischer & Mund
 *     2002) with Neko's conventions, and are pinned by closed forms and dense
 *     brute-force assembly in tests/, not by the reference.
 *
 * Index conventions (shared with the product library, see include/nekoamd.h):
 *   local point  l = i + lx*(j + lx*(k + lx*e)),  i fastest, 0-based
 *   element      e = ex + nx*(ey + ny*ez)
 *   vertex       v = vx + (nx+1)*(vy + (ny+1)*vz)
 *   global id    g = gx + (nx*p+1)*(gy + (ny*p+1)*gz),  gx = ex*p + i, p = lx-1
 *   corner       c = cx + 2*cy + 4*cz of an element
 *   D[i*lx + j]  = d l_j / d r at r = z_i  (row = evaluation point)
 */
#ifndef NEKOTOP_ORACLE_H
#define NEKOTOP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- counter-based RNG (SplitMix64 of seed and index) ------------------ */
uint64_t or_splitmix64(uint64_t seed, uint64_t idx);
double or_u01(uint64_t seed, uint64_t idx); /* [0,1), 53-bit */

/* ---- threads used by the element loops (OpenMP) ------------------------ */
void or_set_threads(int nthreads);
int or_get_threads(void);

/* ---- GLL space ---------------------------------------------------------- */
void or_gll(int lx, double *z, double *w, double *D);

/* ---- box mesh ----------------------------------------------------------- */
void or_box_vertices(int nx, int ny, int nz, double Lx, double Ly, double Lz,
                     double jitter, uint64_t seed, double *vxyz);
void or_element_corners(int nx, int ny, int nz, const double *vxyz,
                        double *corners);
void or_dofmap_gids(int nx, int ny, int nz, int lx, int64_t *gid);
/* faces: bit0 x=0, bit1 x=L, bit2 y=0, bit3 y=L, bit4 z=0, bit5 z=L.
 * Returns the number of local points on those faces; fills out if non-NULL
 * (ascending local index). */
int64_t or_face_mask(int nx, int ny, int nz, int lx, int faces, int32_t *out);
void or_box_rng_uniform(int nx, int ny, int nz, int lx, uint64_t seed,
                        double *out);

/* ---- geometric factors (Neko coef_t) ------------------------------------ */
void or_coef(int lx, int nelv, const double *corners, const double *z,
             const double *w, const double *D, double *x, double *y,
             double *zz, double *G11, double *G12, double *G13, double *G22,
             double *G23, double *G33, double *B, double *jac);

/* ---- ax_helm ------------------------------------------------------------ */
void or_ax_helm(int lx, int nelv, double *w, const double *u, const double *D,
                const double *h1, const double *h2, const double *G11,
                const double *G12, const double *G13, const double *G22,
                const double *G23, const double *G33, const double *B,
                int ifh2);

/* ---- gather-scatter ----------------------------------------------------- */
#define OR_GS_MAXM 64
enum { OR_GS_ADD = 1, OR_GS_MUL = 2, OR_GS_MIN = 3, OR_GS_MAX = 4 };
/* Canonical bucketed map: buckets ordered by multiplicity m ascending; in a
 * bucket, tuples ordered by global id ascending; in a tuple, local indices
 * ascending.  Returns the number of shared occurrences (sum m*cnt). */
int64_t or_gs_build(int64_t n, const int64_t *gid, int32_t *idx, int *nb,
                    int *bm, int64_t *bcnt);
void or_gs_op(double *u, int nb, const int *bm, const int64_t *bcnt,
              const int32_t *idx, int op);

/* ---- SEM bundle used by the solvers ------------------------------------- */
typedef struct {
  int lx, nelv;
  int64_t n;
  const double *D, *G11, *G12, *G13, *G22, *G23, *G33, *B, *mult;
  int gs_nb;
  const int *gs_m;
  const int64_t *gs_cnt;
  const int32_t *gs_idx;
  int64_t nmask;
  const int32_t *mask;
  double volume;
} or_sem_t;

void or_jacobi_update(const or_sem_t *s, const double *h1, const double *h2,
                      int ifh2, double *dinv);
double or_glsc3(int64_t n, const double *a, const double *b, const double *c);
double or_glsc2(int64_t n, const double *a, const double *b);
void or_cg_solve(const or_sem_t *s, const double *h1, const double *h2,
                 int ifh2, const double *dinv, double *x, const double *f,
                 double abstol, int max_iter, int *iter, double *res_start,
                 double *res_final);


/* ---- topology-optimisation kernels (topopt.c) --------------------------- */
void or_brinkman(int64_t n, double *fx, double *fy, double *fz,
                 const double *u, const double *v, const double *w,
                 const double *chi);
void or_update_permeability(int64_t n, double *res, const double *rho,
                            double perm_0, double perm_1, double penalty);
void or_col3_mask(double *a, const double *b, const double *c,
                  const int *mask, int mask_size);
void or_ramp_fwd(int64_t n, double *out, const double *in, double f_min,
                 double f_max, double q, int convex_up);
void or_ramp_bwd(int64_t n, double *df_din, const double *df_dout,
                 const double *in, double f_min, double f_max, double q,
                 int convex_up);
void or_linear_fwd(int64_t n, double *out, const double *in, double f_min,
                   double f_max);
void or_linear_bwd(int64_t n, double *df_din, const double *df_dout,
                   double f_min, double f_max);
void or_grad(int lx, int nelv, const double *D, const double *x,
             const double *y, const double *z, const double *u, double *ux,
             double *uy, double *uz);
void or_dissipation_objective(int lx, int nelv, const double *D,
                              const double *x, const double *y,
                              const double *z, const double *B,
                              const double *u, const double *v,
                              const double *w, const double *chi, double K,
                              double obj_scale, int if_lube, double *out);
void or_dissipation_sensitivity(int64_t n, double *s, const double *u,
                                const double *v, const double *w,
                                const double *ua, const double *va,
                                const double *wa, double K, double obj_scale,
                                int if_lube);
void or_adjoint_lube(int64_t n, double *fx, double *fy, double *fz,
                     const double *u, const double *v, const double *w,
                     const double *chi, double K);
void or_volume_constraint(int64_t n, const double *rho, const double *B,
                          double volume_domain, double v_max, int is_max,
                          double *out);
void or_volume_sensitivity(int64_t n, double *s, double volume_domain,
                           int is_max);
int or_filter_apply(const or_sem_t *s, double r, double abstol, int max_iter,
                    const double *in, double *out);

/* ---- MMA (mma.c) -------------------------------------------------------- */
typedef struct or_mma_s or_mma_t;

void or_adjoint_advection(int lx, int nelv, const double *D, const double *wq,
                          const double *x, const double *y, const double *z,
                          const double *ua, const double *va, const double *wa,
                          const double *ub, const double *vb, const double *wb,
                          double *fx, double *fy, double *fz, double *work);
or_mma_t *or_mma_create(int n, int m, const double *x, double a0, const double *a,
                    const double *c, const double *d, const double *xmin,
                    const double *xmax, int max_iter, double epsimin,
                    double asyinit, double asyincr, double asydecr);
void or_mma_free(or_mma_t *t);
int or_mma_update(or_mma_t *t, int iter, double *x, const double *df0dx,
                  const double *fval, const double *dfdx);
void or_mma_kkt(or_mma_t *t, const double *x, const double *df0dx,
                const double *fval, const double *dfdx);
double or_mma_residumax(const or_mma_t *t);
double or_mma_residunorm(const or_mma_t *t);
void or_mma_get_subproblem(const or_mma_t *t, double *low, double *upp,
                           double *alpha, double *beta, double *p0j,
                           double *q0j, double *pij, double *qij, double *bi);
void or_mma_get_duals(const or_mma_t *t, double *y, double *lambda, double *s,
                      double *mu, double *xsi, double *eta, double *zs);
int or_dgesv(int N, double *A, double *b);

/* ---- Pn/Pn step (pnpn.c) ------------------------------------------------ */
void or_advection(int lx, int nelv, const double *D, const double *x,
                  const double *y, const double *z, const double *B,
                  const double *u, const double *v, const double *w,
                  double *fx, double *fy, double *fz);
void or_curl_local(int lx, int nelv, const double *D, const double *x,
                   const double *y, const double *z, const double *u,
                   const double *v, const double *w, double *cx, double *cy,
                   double *cz);
/* cofactor fields c_b[a] (m = 3b + a), [9][nelv*lx^3] (pnpn.c) */
void or_cofactors(int lx, int nelv, const double *D, const double *x,
                  const double *y, const double *z, double *cof9);
/* dealias.c: Gauss-Legendre rule, Lagrange interpolation and derivative
 * matrices, and the dealiased advection operators (forward, linearised,
 * adjoint), each subtracting from (fx, fy, fz). */
void or_gauss_legendre(int n, double *z, double *w);
void or_lagrange_matrix(int nf, const double *zf, int nt, const double *zt,
                        double *J);
void or_deriv_matrix(int n, const double *z, double *D);
void or_dealias_metric(int lx, int lxd, int nelv, const double *Jm,
                       const double *cof_gll, double *cof);
void or_dealias_advection(int lx, int lxd, int nelv, const double *Jm,
                          const double *Dg, const double *wg,
                          const double *cof, const double *u, const double *v,
                          const double *w, double *fx, double *fy, double *fz);
void or_dealias_linear(int lx, int lxd, int nelv, const double *Jm,
                       const double *Dg, const double *wg, const double *cof,
                       const double *up, const double *vp, const double *wp,
                       const double *ub, const double *vb, const double *wb,
                       double *fx, double *fy, double *fz);
void or_dealias_adjoint(int lx, int lxd, int nelv, const double *Jm,
                        const double *Dg, const double *wg, const double *cof,
                        const double *ua, const double *va, const double *wa,
                        const double *ub, const double *vb, const double *wb,
                        double *fx, double *fy, double *fz);
void or_opgrad(int lx, int nelv, const double *D, const double *wq,
               const double *x, const double *y, const double *z,
               const double *p, double *ox, double *oy, double *oz);
void or_cdtp(int lx, int nelv, const double *D, const double *wq,
             const double *x, const double *y, const double *z,
             const double *tx, const double *ty, const double *tz,
             double *out);
void or_surface_flux(int nx, int ny, int nz, int lx, const double *D,
                     const double *wq, const double *x, const double *y,
                     const double *z, const double *u, const double *v,
                     const double *w, int faces, double *out);
void or_pnpn_coeffs(int tstep, double *bd, double *ext);
void or_pnpn_coeffs3(int tstep, int order, double *bd, double *ext);

#ifdef __cplusplus
}
#endif
#endif
