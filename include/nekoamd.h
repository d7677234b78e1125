/*
 * nekoamd.h -- C-ABI of libnekoamd, the MI355X (gfx950) spectral-element
 * operator core behind neko-top's ax_t / gs_t / ksp_t / pc_t /
 * source_term_t plugin surface.
 *
 * Calling convention (mirrors the reference's own Fortran<->native seam,
 * sources/neko_ext/math/bcknd/device_math_ext.f90:39-105 <->
 * sources/neko_ext/math/bcknd/device/cuda/math_ext.cu:41-109):
 *   - device buffers are passed by value as void*          (type(c_ptr), value)
 *   - scalars are passed by reference                      (real(c_double) :: c)
 *   - every entry point returns an int status, 0 = success; the message of
 *     the last failure is nekoamd_error_string().  (The reference aborts via
 *     CUDA_CHECK, math_ext.cu:56; the Fortran host here turns a non-zero
 *     status into the same kind of abort.)
 *   - all kernels are queued on one in-order stream (the reference's
 *     glb_cmd_queue, math_ext.cu:54); nekoamd_stream() exposes it.
 *
 * Data layout in HBM: every field is nelv contiguous blocks of lx^3 values,
 * i fastest (Neko's x(lx,ly,lz,nelv)); local point l = i + lx*(j + lx*(k +
 * lx*e)), 0-based.  Geometric factors G11,G12,G13,G22,G23,G33 and the mass B
 * are separate fields (Neko's coef_t), FP64.
 *
 * Index conventions: device index lists (Dirichlet masks, gs maps) are
 * 0-based int32.  The hip_*_mask entry points keep the reference's 1-based
 * Fortran masks (math_ext_kernel.h:49 subtracts 1) for drop-in use.
 */
#ifndef NEKOAMD_H
#define NEKOAMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NEKOAMD_LX_MAX 10
#define NEKOAMD_GS_MAX_MULT 64

enum {
  NEKOAMD_HOST_TO_DEVICE = 1,
  NEKOAMD_DEVICE_TO_HOST = 2,
  NEKOAMD_DEVICE_TO_DEVICE = 3
};

/* Neko's GS_OP_ADD / GS_OP_MUL / GS_OP_MIN / GS_OP_MAX */
enum { NEKOAMD_GS_ADD = 1, NEKOAMD_GS_MUL = 2, NEKOAMD_GS_MIN = 3, NEKOAMD_GS_MAX = 4 };

/* Box faces, for the synthetic channel's boundary zones. */
enum {
  NEKOAMD_FACE_XMIN = 1,  /* inflow 'v'  */
  NEKOAMD_FACE_XMAX = 2,  /* outflow 'o' */
  NEKOAMD_FACE_YMIN = 4,
  NEKOAMD_FACE_YMAX = 8,
  NEKOAMD_FACE_ZMIN = 16,
  NEKOAMD_FACE_ZMAX = 32
};

/* ======================================================================
 * Device layer.  Replaces Neko's device_map / device_memcpy / device_free
 * as called at sources/topology_optimization/design.f90:150-155, :189-191,
 * :214 and sources/adjoint/adjoint_pnpn.f90:441-474.
 * ==================================================================== */
const char* nekoamd_error_string(void);
int nekoamd_device_init(const int* device);
int nekoamd_device_finalize(void);
int nekoamd_device_name(char* buf, const int* len);
void* nekoamd_stream(void);
int nekoamd_device_alloc(void** ptr, const int64_t* bytes);
int nekoamd_device_free(void* ptr);
int nekoamd_host_alloc_pinned(void** ptr, const int64_t* bytes);
int nekoamd_host_free_pinned(void* ptr);
int nekoamd_device_memcpy(void* dst, const void* src, const int64_t* bytes,
                          const int* kind, const int* sync);
int nekoamd_device_sync(void);
int nekoamd_event_create(void** ev);
int nekoamd_event_destroy(void* ev);
int nekoamd_event_record(void* ev);
int nekoamd_event_sync(void* ev);
int nekoamd_event_elapsed(void* ev0, void* ev1, float* ms);

/* ======================================================================
 * Space, mesh and dofmap (host).  What Neko's space_t / dofmap_t / mesh
 * set-up provides before sources/adjoint/adjoint_scheme.f90:331-343
 * (Xh, dm_Xh, gs_Xh%init, c_Xh%init).
 * ==================================================================== */
uint64_t nekoamd_rng_u64(const uint64_t* seed, const uint64_t* idx);
/* GLL nodes z, weights w, derivative matrix dx[i*lx+j] = l_j'(z_i). */
int nekoamd_gll(const int* lx, double* z, double* w, double* dx);
/* Vertices of an nx*ny*nz box [0,L0]x[0,L1]x[0,L2] (lxyz[3]); interior
 * coordinates jittered by U(-jitter,jitter)*h per axis from SplitMix64(seed,
 * 3*v+axis).  vxyz has 3*(nx+1)*(ny+1)*(nz+1) entries. */
int nekoamd_box_vertices(const int* nx, const int* ny, const int* nz,
                         const double* lxyz, const double* jitter,
                         const uint64_t* seed, double* vxyz);
/* Element corners (nelv*8*3), corner c = cx + 2cy + 4cz. */
int nekoamd_box_corners(const int* nx, const int* ny, const int* nz,
                        const double* vxyz, double* corners);
/* Global ids of all nelv*lx^3 local points (lattice numbering). */
int nekoamd_box_gids(const int* nx, const int* ny, const int* nz,
                     const int* lx, int64_t* gid);
/* Local indices (0-based, ascending) of the points on the given faces;
 * out may be NULL to count only. */
int nekoamd_box_face_mask(const int* nx, const int* ny, const int* nz,
                          const int* lx, const int* faces, int32_t* out,
                          int64_t* count);

/* ======================================================================
 * Geometric factors on the device.  Replaces Neko coef_t%init (G11..G33,
 * B, jac; used by every ax_t%compute call, e.g.
 * sources/adjoint/adjoint_pnpn.f90:716-723, :734-736).
 * corners_d: nelv*8*3 FP64; z_d, w_d: lx; dx_d: lx*lx.  Outputs are FP64
 * fields; any output pointer may be NULL except G*, B.
 * ==================================================================== */
int nekoamd_coef_geo(void* x_d, void* y_d, void* z_d, void* g11, void* g12,
                     void* g13, void* g22, void* g23, void* g33, void* B,
                     void* jac, const void* corners_d, const void* zgll_d,
                     const void* wgll_d, const void* dx_d, const int* lx,
                     const int* nelv);

/* ======================================================================
 * ax_helm: w = sum_a D_a^T (h1 sum_b G_ab D_b u) [+ h2 B u if *ifh2].
 * Replaces Neko's ax_helm_t%compute(w, u, coef, msh, Xh), built by
 * ax_helm_factory at sources/adjoint/adjoint_pnpn.f90:207 (velocity), :217
 * (pressure) and sources/mapping_functions/PDE_filter_mapping.f90:177.
 * h2 and B may be NULL when *ifh2 == 0.
 *   _f64      : FP64 u, w, h1, h2; FP64 G, B
 *   _f32      : FP32 u, w, h1, h2; FP64 G, B (geometric factors stay FP64)
 *   _f32g32   : FP32 everything
 * ==================================================================== */
int nekoamd_ax_helm_f64(void* w, const void* u, const void* dx, const void* h1,
                        const void* h2, const void* g11, const void* g12,
                        const void* g13, const void* g22, const void* g23,
                        const void* g33, const void* B, const int* lx,
                        const int* nelv, const int* ifh2);
/* Three fields through the same FP64 operator in one launch: w1 = A u1,
 * w2 = A u2, w3 = A u3, each equal bit for bit to nekoamd_ax_helm_f64 on that
 * field; the operator data are read once.  Replaces Neko's
 * ax_t%compute_vector(au, av, aw, u, v, w, coef, msh, Xh), the call of the
 * velocity residual (Neko's pnpn_vel_res, called at
 * sources/adjoint/adjoint_pnpn.f90:747-753). */
int nekoamd_ax_helm_vector_f64(void* w1, void* w2, void* w3, const void* u1,
                               const void* u2, const void* u3, const void* dx,
                               const void* h1, const void* h2,
                               const void* g11, const void* g12,
                               const void* g13, const void* g22,
                               const void* g23, const void* g33,
                               const void* B, const int* lx, const int* nelv,
                               const int* ifh2);
/* The velocity residual of the three components in one launch:
 * r_c = (F_c - g_c) - A u_c with the FP64 operator, i.e.
 * nekoamd_ax_helm_vector_f64 into r followed by nekoamd_pnpn_vel_res(r_c,
 * F_c, g_c) (Neko's pnpn_vel_res, sources/adjoint/adjoint_pnpn.f90:747-753);
 * the same values. */
int nekoamd_pnpn_vel_res_vector(void* r1, void* r2, void* r3, const void* u1,
                                const void* u2, const void* u3,
                                const void* F1, const void* F2,
                                const void* F3, const void* g1,
                                const void* g2, const void* g3,
                                const void* dx, const void* h1,
                                const void* h2, const void* g11,
                                const void* g12, const void* g13,
                                const void* g22, const void* g23,
                                const void* g33, const void* B, const int* lx,
                                const int* nelv, const int* ifh2);
int nekoamd_ax_helm_f32(void* w, const void* u, const void* dx, const void* h1,
                        const void* h2, const void* g11, const void* g12,
                        const void* g13, const void* g22, const void* g23,
                        const void* g33, const void* B, const int* lx,
                        const int* nelv, const int* ifh2);
int nekoamd_ax_helm_f32g32(void* w, const void* u, const void* dx,
                           const void* h1, const void* h2, const void* g11,
                           const void* g12, const void* g13, const void* g22,
                           const void* g23, const void* g33, const void* B,
                           const int* lx, const int* nelv, const int* ifh2);

/* ======================================================================
 * gs_op.  Replaces Neko gs_t: init(dofmap) at
 * sources/adjoint/adjoint_scheme.f90:341, op(field, GS_OP_ADD) at
 * sources/adjoint/adjoint_pnpn.f90:725, :755-757 and
 * sources/mapping_functions/PDE_filter_mapping.f90:249, :338, free at
 * adjoint_scheme.f90:913.
 * The map is built on the host from the n global ids: buckets of equal
 * multiplicity m (ascending); tuples by global id; members by local index.
 * Values of a tuple are combined in member order, so results are
 * bit-identical to a sequential CPU gather-scatter with the same map.
 * ==================================================================== */
int nekoamd_gs_init(void** gs, const int64_t* gid, const int64_t* n);
int nekoamd_gs_free(void* gs);
int nekoamd_gs_op_f64(void* gs, void* u, const int* op);
int nekoamd_gs_op_f32(void* gs, void* u, const int* op);
/* u = gs(B u) Binv (Binv = 1 / gs(B)), FP64, in one pass over the shared
 * points: the mass average of the curl-curl pressure residual (Neko's
 * col2(B), gs_op ADD, col2(Binv) inside the residual the step computes
 * before the pressure solve, sources/adjoint/adjoint_pnpn.f90:716-723).
 * Shared points get the same products and sum in the same order as the
 * three passes; other points keep u exactly.  B and Binv are gathered into
 * the map's order on the first call with a given pair of arrays and kept
 * (keyed by their addresses: the mass of a mesh is fixed once built). */
int nekoamd_gs_mass_average(void* gs, void* u, const void* B,
                            const void* Binv);
/* nb buckets; bm[nb], bcnt[nb] (arrays of NEKOAMD_GS_MAX_MULT); nocc. */
int nekoamd_gs_map_info(void* gs, int* nb, int* bm, int64_t* bcnt,
                        int64_t* nocc);
/* Copy the device index map (nocc int32) to the host. */
int nekoamd_gs_map_copy(void* gs, int32_t* idx);

/* ======================================================================
 * Vector kernels (Neko device_math as used on the path, FP64).
 * ==================================================================== */
int nekoamd_copy(void* a, const void* b, const int64_t* n);
int nekoamd_rzero(void* a, const int64_t* n);
int nekoamd_cfill(void* a, const double* c, const int64_t* n);
int nekoamd_cmult(void* a, const double* c, const int64_t* n);
int nekoamd_add2(void* a, const void* b, const int64_t* n);
int nekoamd_add2s2(void* a, const void* b, const double* c, const int64_t* n);
/* a = b + c (Neko cadd2) */
int nekoamd_cadd2(void* a, const void* b, const double* c, const int64_t* n);
int nekoamd_col2(void* a, const void* b, const int64_t* n);
int nekoamd_col3(void* a, const void* b, const void* c, const int64_t* n);
int nekoamd_addcol3(void* a, const void* b, const void* c, const int64_t* n);
int nekoamd_subcol3(void* a, const void* b, const void* c, const int64_t* n);
int nekoamd_invcol1(void* a, const int64_t* n);
int nekoamd_f64_to_f32(void* a32, const void* b64, const int64_t* n);
int nekoamd_f32_to_f64(void* a64, const void* b32, const int64_t* n);
/* Zero u at the 0-based local indices mask[0..nmask) (Dirichlet bc apply,
 * Neko bc_list_apply_scalar, sources/adjoint/adjoint_pnpn.f90:726). */
int nekoamd_bc_zero(void* u, const void* mask, const int64_t* nmask);
int nekoamd_bc_zero_f32(void* u, const void* mask, const int64_t* nmask);
/* Deterministic global sums (fixed reduction tree), result on the host. */
int nekoamd_glsc2(double* res, const void* a, const void* b, const int64_t* n);
int nekoamd_glsc3(double* res, const void* a, const void* b, const void* c,
                  const int64_t* n);

/* Masked kernels of the reference (1-based Fortran masks, like
 * sources/neko_ext/math/bcknd/device/cuda/math_ext_kernel.h:42-114 and the
 * wrappers at math_ext.cu:49-108; glsc2_mask follows math_ext.f90:99-118). */
int hip_cadd_mask(void* a, const double* c, const int* size, const void* mask,
                  const int* mask_size);
int hip_invcol1_mask(void* a, const int* size, const void* mask,
                     const int* mask_size);
int hip_col2_mask(void* a, const void* b, const int* size, const void* mask,
                  const int* mask_size);
int hip_col3_mask(void* a, const void* b, const void* c, const int* size,
                  const void* mask, const int* mask_size);
int hip_sub3_mask(void* a, const void* b, const void* c, const int* size,
                  const void* mask, const int* mask_size);
int hip_glsc2_mask(double* res, const void* a, const void* b, const int* size,
                   const void* mask, const int* mask_size);

/* ======================================================================
 * Jacobi preconditioner.  Replaces Neko jacobi_t / device_jacobi_t:
 * init at sources/adjoint/adjoint_scheme.f90:1068-1076, update() at
 * sources/adjoint/adjoint_pnpn.f90:732, :781, solve(z, r, n) inside CG.
 * update: d = 1 / gs_ADD( h1 * sum_l (D_li^2 G11 + D_lj^2 G22 + D_lk^2 G33)
 *                         [+ h2 B] ).
 * ==================================================================== */
int nekoamd_jacobi_update(void* d, void* gs, const void* dx, const void* g11,
                          const void* g22, const void* g33, const void* h1,
                          const void* h2, const void* B, const int* lx,
                          const int* nelv, const int* ifh2);
int nekoamd_jacobi_solve(void* z, const void* r, const void* d,
                         const int64_t* n);

/* ======================================================================
 * Preconditioned CG with device-resident scalars.  Replaces Neko's
 * ksp_t%solve for type 'cg' (krylov_solver_factory at
 * sources/adjoint/adjoint_scheme.f90:1046-1055; solve at
 * sources/adjoint/adjoint_pnpn.f90:734-736, monitor logged at :812).
 * alpha, beta, the residual norm and the stop flag never leave the device;
 * the host only polls the stop flag every *poll iterations.  The state owns
 * its work vectors.
 * Norm: ||r|| = sqrt(sum r*r*mult) / sqrt(volume); x0 = 0; stop after the
 * first iteration with ||r|| < abstol, or after max_iter iterations.
 *
 * NEKOAMD_CG_FP64:  every vector and the operator in FP64.
 * NEKOAMD_CG_MIXED: r, p, w, Jacobi, mult, h1, h2, the geometric factors
 *   and B stored in FP32 (x in FP64, the correction since the last reliable
 *   update accumulated in FP32 and folded into x there; the element
 *   contraction and every inner product in FP64 arithmetic).  The FP32 copies of G11..G33, B and mult
 *   are made on the first solve with a given set of those arrays and kept,
 *   keyed by their addresses: a caller that rewrites them in place (a
 *   moving mesh) creates a new state.  FP64 reliable updates: r = f - A x
 *   with the FP64 operator whenever the iterated residual has dropped 10x
 *   since the last update, and before stopping.  The solve stops on the
 *   FP64 residual only (res_final is that residual).  No reference
 *   counterpart: the FP32 iteration of the mixed-precision configuration.
 * ==================================================================== */
#define NEKOAMD_CG_FP64 0
#define NEKOAMD_CG_MIXED 1

typedef struct {
  int iter;        /* iterations (FP32 iterations in mixed mode) */
  int done;        /* 1: converged or max_iter reached */
  double res_start;
  double res_final;
  int nupdate;     /* reliable updates (mixed mode) */
  int pad;
} nekoamd_cg_monitor_t;

/* n = nelv * lx^3 points; precision = NEKOAMD_CG_FP64 or NEKOAMD_CG_MIXED */
int nekoamd_cg_create(void** cg, const int64_t* n, const int* nelv,
                      const int* precision);
int nekoamd_cg_free(void* cg);
/* x = A^-1 f for A = QQ^T (h1 A_L + h2 B) with the Dirichlet mask (0-based
 * indices mask[0..nmask)), preconditioned by the Jacobi diagonal d
 * (nekoamd_jacobi_update).  f must be continuous (gs-assembled) and zero on
 * the mask, as the reference's callers guarantee (adjoint_pnpn.f90:725-726).
 * Blocks until the solve is done; fills mon. */
int nekoamd_cg_solve(void* cg, void* x, const void* f, const void* d,
                     const void* mult, const double* volume,
                     const double* abstol, const int* max_iter,
                     const int* poll, void* gs, const void* mask,
                     const int64_t* nmask, const void* dx, const void* h1,
                     const void* h2, const void* g11, const void* g12,
                     const void* g13, const void* g22, const void* g23,
                     const void* g33, const void* B, const int* lx,
                     const int* ifh2, nekoamd_cg_monitor_t* mon);
/* Flexible PCG (Polak-Ribiere beta, Notay 2000) for the pressure solve
 * (ksp_prs%solve, sources/adjoint/adjoint_pnpn.f90:734-736): Jacobi, plus,
 * when mlp is not null, the element network's learned correction every
 * `every`-th iteration -- z = d r + gamma q with q = mask(mult QQ^T N(r))
 * and its Galerkin weight gamma = <q,r>/<q,Aq> (0 if <q,Aq> <= 0).  The
 * reference has no learned component; with mlp = null the iterates are
 * those of nekoamd_cg_solve's FP64 PCG.  Operator, mask and monitor as
 * nekoamd_cg_solve; the network (nekoamd_mlp_create) must be made for the
 * same mesh, lx = 8. */
int nekoamd_fcg_create(void** st, const int64_t* n, const int* nelv);
int nekoamd_fcg_free(void* st);
int nekoamd_fcg_solve(void* st, void* x, const void* f, const void* d,
                      const void* mult, const double* volume,
                      const double* abstol, const int* max_iter,
                      const int* poll, void* gs, const void* mask,
                      const int64_t* nmask, const void* dx, const void* h1,
                      const void* h2, const void* g11, const void* g12,
                      const void* g13, const void* g22, const void* g23,
                      const void* g33, const void* B, const int* lx,
                      const int* ifh2, void* mlp, const int* every,
                      nekoamd_cg_monitor_t* mon);
/* solve_coupled (sources/adjoint/adjoint_pnpn.f90:784-786): the three
 * velocity systems A x_c = f_c with one operator, Jacobi diagonal and mask,
 * on three CG states made with the same n, nelv and precision.  In mixed
 * mode the three solves share one set-up (FP32 operator data and Jacobi
 * diagonal, kept in the first state) and one driver (their iterations go in
 * the same batches, one host read per batch for the three); each component
 * keeps its own scalars, stopping test and reliable updates (mon[0..2]).
 * FP64 states (or lx outside [4, 10]) run the three solves one after the
 * other. */
int nekoamd_cg_solve3(void* const* cg, void* const* x, const void* const* f,
                      const void* d, const void* mult, const double* volume,
                      const double* abstol, const int* max_iter,
                      const int* poll, void* gs, const void* mask,
                      const int64_t* nmask, const void* dx, const void* h1,
                      const void* h2, const void* g11, const void* g12,
                      const void* g13, const void* g22, const void* g23,
                      const void* g33, const void* B, const int* lx,
                      const int* ifh2, nekoamd_cg_monitor_t* mon);
/* The monitor of the last solve. */
int nekoamd_cg_status(void* cg, nekoamd_cg_monitor_t* mon);

/* ---- GMRES (Neko gmres_t, krylov_solver_factory 'gmres') -----------------
 * Restarted flexible GMRES with restart length m (Neko's default 30), the
 * pressure solver of the shipped cases (rugby_ball.case:46-52), solving
 * where the reference solves (adjoint_pnpn.f90:734-736).  x0 = 0; classical
 * Gram-Schmidt with <.,.>_mult; stops when |gam_{j+1}| / sqrt(volume) <
 * abstol or after max_iter iterations.  The preconditioner is
 * NEKOAMD_PC_IDENT (pc unused), NEKOAMD_PC_JACOBI (pc = the inverse diagonal
 * of nekoamd_jacobi_update) or NEKOAMD_PC_HSMG (pc = nekoamd_hsmg_create
 * handle).  Same operator arguments as nekoamd_cg_solve; the monitor reports
 * iter, done, res_start, res_final. */
#define NEKOAMD_PC_IDENT 0
#define NEKOAMD_PC_JACOBI 1
#define NEKOAMD_PC_HSMG 2
#define NEKOAMD_GMRES_MAX_M 60
int nekoamd_gmres_create(void** st, const int64_t* n, const int* m);
int nekoamd_gmres_free(void* st);
int nekoamd_gmres_solve(void* st, void* x, const void* f, const int* pc_kind,
                        void* pc, const void* mult, const double* volume,
                        const double* abstol, const int* max_iter, void* gs,
                        const void* mask, const int64_t* nmask, const void* dx,
                        const void* h1, const void* h2, const void* g11,
                        const void* g12, const void* g13, const void* g22,
                        const void* g23, const void* g33, const void* B,
                        const int* lx, const int* ifh2,
                        nekoamd_cg_monitor_t* mon);
int nekoamd_gmres_status(void* st, nekoamd_cg_monitor_t* mon);

/* ---- hybrid Schwarz multigrid (Neko pc_hsmg, precon_factory 'hsmg') ------
 * The pressure preconditioner of the shipped cases (rugby_ball.case:46-52;
 * adjoint_scheme.f90:1077-1087 builds it).  Levels lx, 4 (lx >= 5; 3 for
 * lx = 4) and 2 of the box mesh with the given element corners (host,
 * nelv*8*3, corner c = cx + 2cy + 4cz); overlapping Schwarz with fast-
 * diagonalisation local solves on the two finer levels, crs_iter PCG
 * iterations on the coarse one (Neko: 10); dirichlet_faces = the box faces
 * (NEKOAMD_FACE_*) that are Dirichlet for the preconditioned operator.
 * lx must be 4, 6 or 8.  precision: NEKOAMD_CG_FP64, or NEKOAMD_CG_MIXED for
 * an FP32 data path (Schwarz, FDM, transfers, fine and mid gs; the coarse PCG
 * stays FP64), an inexact preconditioner for the FP64 flexible GMRES.
 * apply: z = M^-1 r (fine-level FP64 fields). */
int nekoamd_hsmg_create(void** st, const int* nx, const int* ny, const int* nz,
                        const int* lx, const double* corners,
                        const int* dirichlet_faces, const int* crs_iter,
                        const int* precision);
int nekoamd_hsmg_free(void* st);
int nekoamd_hsmg_apply(void* st, void* z, const void* r);


/* ======================================================================
 * Solution projection for successive right-hand sides (Fischer 1998).
 * Replaces Neko's projection_t as the reference drives it: proj_prs and
 * proj_u/v/w (sources/adjoint/adjoint_pnpn.f90:102-105), init(n, L,
 * activ_step) (:365-373), pre_solving (:729-730, :777-779) and
 * post_solving (:739-740, :788-793).  The basis (up to L previous
 * solutions, A-orthogonal, with their images A x_i and squared A-norms)
 * stays in HBM.
 * Inner products are mult-weighted.
 * ==================================================================== */
#define NEKOAMD_PROJ_MAX_L 64

int nekoamd_proj_create(void** proj, const int64_t* n, const int* L);
int nekoamd_proj_free(void* proj);
/* Forget the basis (the reference does this when dt changes). */
int nekoamd_proj_clear(void* proj);
/* Current basis size m. */
int nekoamd_proj_size(void* proj, int* m);
/* Solutions kept when the full basis restarts, 0 <= keep < L (default 0:
 * Neko's restart from the newest solution alone).  keep > 0 restarts from
 * the span of the keep newest solutions, formed from the stored basis in
 * one streaming pass.  No reference counterpart (the restart policy of
 * Neko's projection_t is keep = 0). */
int nekoamd_proj_set_keep(void* proj, const int* keep);
/* FP32 basis (*on != 0; the basis must be empty): the previous solutions
 * are stored in FP32 and their images under the operator are not stored;
 * the arithmetic stays FP64.  pre then forms xbar from the FP32 basis and
 * b -= A xbar with one operator application (the operator of the last
 * post), so the projected right-hand side is exact for the xbar it hands
 * back; post forms A of the new vector once.  About half the bytes per
 * solve of the FP64 basis, for a guess that differs from it by the FP32
 * rounding of the basis.  No reference counterpart (Neko's projection_t
 * keeps an FP64 basis and its images). */
int nekoamd_proj_set_fp32(void* proj, const int* on);
/* pre_solving: alpha_i = <b, x_i>; xbar = sum alpha_i x_i;
 * b -= sum alpha_i A x_i.  No-op while the basis is empty. */
int nekoamd_proj_pre(void* proj, void* b, const void* mult);
/* One extra direction q with its image aq = mask(QQ^T A q), Galerkin-scaled:
 * gamma = <q,b>/<q,aq> (0 unless <q,aq> > 0); xbar += gamma q;
 * b -= gamma aq.  Used for the learned initial guess.  No reference
 * counterpart. */
int nekoamd_proj_augment(void* proj, void* b, const void* q, const void* aq,
                         const void* mult, double* gamma);
/* post_solving: x += xbar; x joins the basis (restarting it when full)
 * unless it is A-dependent on it; its image is mask(QQ^T A x) with the
 * given operator (the ax_helm arguments). */
int nekoamd_proj_post(void* proj, void* x, const void* mult, void* gs,
                      const void* mask, const int64_t* nmask, const void* dx,
                      const void* h1, const void* h2, const void* g11,
                      const void* g12, const void* g13, const void* g22,
                      const void* g23, const void* g33, const void* B,
                      const int* lx, const int* nelv, const int* ifh2);


/* ======================================================================
 * Learned solver aid: an element-local network (518 -> 256 -> 256 -> 512,
 * SiLU, BF16 MFMA with FP32 accumulation) that adds one Galerkin-scaled
 * direction to the initial guess of a solve, trained on the device on
 * (residual, solution) pairs harvested from the solver.  No reference
 * counterpart: the reference's initial-guess accelerator is the projection
 * above (adjoint_pnpn.f90:102-105, :729-741).  Elements of lx = 8 only.
 * Inputs per element: the 512 residual values and the means over its 6 face
 * neighbours (x-fastest box numbering), divided by the RMS of the residual.
 * Parameters (FP32): W1[256][518] b1[256] W2[256][256] b2[256] W3[512][256]
 * b3[512], row-major.
 * ==================================================================== */
#define NEKOAMD_MLP_NPARAM 330240
#define NEKOAMD_MLP_KIN 544   /* padded input width of a feature row */

/* capacity: harvested pairs kept (ring); batch: training rows per step
 * (multiple of 32); seed: weight initialisation. */
int nekoamd_mlp_create(void** mlp, const int* nx, const int* ny, const int* nz,
                       const int* lx, const int* capacity, const int* batch,
                       const int64_t* seed);
int nekoamd_mlp_free(void* mlp);
int nekoamd_mlp_get_params(void* mlp, float* host_params);
int nekoamd_mlp_set_params(void* mlp, const float* host_params);
/* Y[M][512] (FP32) = network(X), X[M][544] BF16 feature rows (device);
 * H1, H2 (M x 256 BF16, optional): the hidden activations. */
int nekoamd_mlp_forward(void* mlp, const void* X, const int* M, void* Y,
                        void* H1, void* H2);
/* loss = mean (Y - T)^2 over M x 512, g (device, NEKOAMD_MLP_NPARAM FP32,
 * the parameter layout) = its gradient; parameters unchanged. */
int nekoamd_mlp_loss_grad(void* mlp, const void* X, const void* T,
                          const int* M, double* loss, void* g);
/* Store one (residual b, solution x) pair of fields (FP64, n = nelv*512). */
int nekoamd_mlp_harvest(void* mlp, const void* b, const void* x);
int nekoamd_mlp_pairs(void* mlp, int* npairs);
/* Adam (lr, beta 0.9 / 0.999) over the stored pairs, targets x / RMS(b)
 * normalised by their RMS; mean loss of the first and last epoch. */
int nekoamd_mlp_train(void* mlp, const int* epochs, const int* batch,
                      const float* lr, const int64_t* seed, double* loss_first,
                      double* loss_last);
/* Before a solve of A x = b: keep b for harvesting (save) and/or apply the
 * guess (apply): q = mask(mult QQ^T N(b)), gamma = <q,b>/<q,Aq> (0 unless
 * <q,Aq> > 0), b -= gamma A q.  The operator arguments are ax_helm's. */
int nekoamd_mlp_pre(void* mlp, void* b, const void* mult, void* gs,
                    const void* mask, const int64_t* nmask, const void* dx,
                    const void* h1, const void* h2, const void* g11,
                    const void* g12, const void* g13, const void* g22,
                    const void* g23, const void* g33, const void* B,
                    const int* lx, const int* ifh2, const int* apply,
                    const int* save, double* gamma);
/* After the solve: x += gamma q if a guess was applied; then, if harvest
 * and b was kept, store the pair (b, x). */
int nekoamd_mlp_post(void* mlp, void* x, const int* harvest);


/* ======================================================================
 * Topology-optimisation kernels (FP64, fused single passes).
 * ==================================================================== */
/* f -= chi u per component: simple_brinkman_source_term_t%compute_,
 * sources/source_terms/simple_brinkman_source_term.f90:139-153. */
int nekoamd_brinkman(void* fx, void* fy, void* fz, const void* u,
                     const void* v, const void* w, const void* chi,
                     const int64_t* n);
/* f -= B chi u per component: the Brinkman source term of
 * simple_brinkman_source_term.f90:139-153 with the mass matrix the step
 * applies to the source terms after them (opcolv, adjoint_pnpn.f90:672-676)
 * fused in; the explicit Brinkman mode of the Pn/Pn step. */
int nekoamd_brinkman_mass(void* fx, void* fy, void* fz, const void* u,
                          const void* v, const void* w, const void* chi,
                          const void* B, const int64_t* n);
/* resistance = -(k1 + (k0-k1)(q+1) rho/(q+rho)): design_t%update_permeability,
 * sources/topology_optimization/design.f90:316-348 (the masked force itself
 * is hip_col3_mask, design.f90:266-294). */
int nekoamd_update_permeability(void* res, const void* rho,
                                const double* perm_0, const double* perm_1,
                                const double* penalty, const int64_t* n);
/* RAMP_mapping_t%apply_forward / apply_backward,
 * sources/mapping_functions/RAMP_mapping.f90:138-267 (the backward is the
 * reference's formula, including its convex-up form). */
int nekoamd_ramp_fwd(void* out, const void* in, const double* f_min,
                     const double* f_max, const double* q,
                     const int* convex_up, const int64_t* n);
int nekoamd_ramp_bwd(void* df_din, const void* df_dout, const void* in,
                     const double* f_min, const double* f_max, const double* q,
                     const int* convex_up, const int64_t* n);
/* linear_mapping_t, sources/mapping_functions/linear_mapping.f90:104-136. */
int nekoamd_linear_fwd(void* out, const void* in, const double* f_min,
                       const double* f_max, const int64_t* n);
int nekoamd_linear_bwd(void* df_din, const void* df_dout, const double* f_min,
                       const double* f_max, const int64_t* n);
/* dF/dchi = -(u.u_adj) [+ K obj_scale |u|^2]:
 * minimum_dissipation_objective_function_compute_sensitivity,
 * sources/objectives/minimum_dissipation_objective_function.f90:260-301. */
int nekoamd_dissipation_sensitivity(void* s, const void* u, const void* v,
                                    const void* w, const void* ua,
                                    const void* va, const void* wa,
                                    const double* K, const double* obj_scale,
                                    const int* if_lube, const int64_t* n);
/* f_adj += K chi u: adjoint_lube_source_term_compute,
 * sources/source_terms/adjoint_lube_source_term.f90:173-206. */
int nekoamd_adjoint_lube(void* fx, void* fy, void* fz, const void* u,
                         const void* v, const void* w, const void* chi,
                         const double* K, const int64_t* n);
/* Element-local gradient (Neko grad, used by the objective at
 * minimum_dissipation_objective_function.f90:200-213); lx <= 8. */
int nekoamd_grad(void* ux, void* uy, void* uz, const void* u, const void* x,
                 const void* y, const void* z, const void* dx, const int* lx,
                 const int* nelv);
/* minimum_dissipation_objective_function_compute,
 * minimum_dissipation_objective_function.f90:186-254: out[0] = F,
 * out[1] = dissipation, out[2] = lube value (sum chi (u+v+w) B). */
int nekoamd_dissipation_objective(double* out, const void* u, const void* v,
                                  const void* w, const void* chi,
                                  const void* B, const void* x, const void* y,
                                  const void* z, const void* dx, const int* lx,
                                  const int* nelv, const double* K,
                                  const double* obj_scale, const int* if_lube);
/* No counterpart in the reference: the objective above for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and results; the
 * per-element partials live in a static device array of the library
 * (4 MiB per device, taken at nekoamd_device_init) instead of a per-call
 * hipMalloc, and the two sums come back through a stack buffer instead of
 * pinned memory.  nelv may be at most 262144 (larger sizes return an
 * argument error before any device work).  Calls that overlap on one
 * device (several host threads) race on that array. */
int nekoamd_dissipation_objective_noalloc(
    double* out, const void* u, const void* v, const void* w, const void* chi,
    const void* B, const void* x, const void* y, const void* z,
    const void* dx, const int* lx, const int* nelv, const double* K,
    const double* obj_scale, const int* if_lube);
/* volume_constraint_compute / _compute_sensitivity,
 * sources/objectives/volume_constraint.f90:177-242: out[0] = constraint
 * value, out[1] = volume fraction. */
int nekoamd_volume_constraint(double* out, const void* rho, const void* B,
                              const double* volume_domain, const double* v_max,
                              const int* is_max, const int64_t* n);
int nekoamd_volume_sensitivity(void* s, const double* volume_domain,
                               const int* is_max, const int64_t* n);


/* ======================================================================
 * MMA (method of moving asymptotes) with the n design variables in HBM.
 * Replaces mma_t (sources/mma/mma.f90:49-458): init at mma.f90:132-184
 * (epsimin <= 0 selects the default 1e-9 sqrt(m+n)), update(iter, x, df0dx,
 * fval, dfdx) at mma.f90:297-324 = gensub + dual-primal subsolve
 * (sources/mma/bcknd/cpu/mma_cpu.f90:8-579), KKT at mma_cpu.f90:581-641,
 * get_residumax / get_residunorm.  x, xmin, xmax, df0dx: device (n);
 * dfdx: device, m x n column-major; a, c, d, fval: host (m); 1 <= m <= 4.
 * A singular Newton system (the reference's DGESV failure, which stops the
 * program, mma_cpu.f90:440-446) returns status 5.
 * ==================================================================== */
int nekoamd_mma_create(void** mma, const int* n, const int* m, const void* x,
                       const double* a0, const double* a, const double* c,
                       const double* d, const void* xmin, const void* xmax,
                       const int* max_iter, const double* epsimin,
                       const double* asyinit, const double* asyincr,
                       const double* asydecr);
int nekoamd_mma_free(void* mma);
int nekoamd_mma_update(void* mma, const int* iter, void* x, const void* df0dx,
                       const double* fval, const void* dfdx);
int nekoamd_mma_kkt(void* mma, const void* x, const void* df0dx,
                    const double* fval, const void* dfdx);
/* res[0] = residumax, res[1] = residunorm, res[2] = Newton iterations of
 * the last update. */
int nekoamd_mma_residuals(void* mma, double* res);
int nekoamd_mma_duals(void* mma, double* y, double* lambda, double* s,
                      double* mu, double* zs);

/* ======================================================================
 * Pn/Pn step operators (the calls of sources/adjoint/adjoint_pnpn.f90:
 * 619-827 into Neko's sumab, adv, makeabf/makebdf, pnpn_prs_res and
 * pnpn_vel_res).  Element kernels need x, y, z (coef_geo) and lx <= 8.
 * ==================================================================== */
/* B f = -B (u.grad) u  (adv%compute, adjoint_pnpn.f90:680-682) */
int nekoamd_advection(void* fx, void* fy, void* fz, const void* u,
                      const void* v, const void* w, const void* B,
                      const void* x, const void* y, const void* z,
                      const void* dx, const int* lx, const int* nelv);
/* element-local curl (Neko curl before its mass averaging) */
int nekoamd_curl(void* cx, void* cy, void* cz, const void* u, const void* v,
                 const void* w, const void* x, const void* y, const void* z,
                 const void* dx, const int* lx, const int* nelv);
/* B grad p (Neko opgrad, used by pnpn_vel_res) */
int nekoamd_opgrad(void* ox, void* oy, void* oz, const void* p, const void* x,
                   const void* y, const void* z, const void* dx,
                   const void* wgll, const int* lx, const int* nelv);
/* sum over components of Neko cdtp (weak divergence, pnpn_prs_res) */
int nekoamd_cdtp(void* out, const void* tx, const void* ty, const void* tz,
                 const void* x, const void* y, const void* z, const void* dx,
                 const void* wgll, const int* lx, const int* nelv);
/* The divergence part of the pressure residual in one pass:
 * out = cdtp(binv t) - b - s c, i.e. col2(t_a, Binv) on the three fields,
 * nekoamd_cdtp, then nekoamd_sub2s(out, b, c, s) (Neko's pnpn_prs_res,
 * called at sources/adjoint/adjoint_pnpn.f90:716-723); t is not modified. */
int nekoamd_pnpn_prs_div(void* out, const void* tx, const void* ty,
                         const void* tz, const void* binv, const void* b,
                         const void* c, const double* s, const void* x,
                         const void* y, const void* z, const void* dx,
                         const void* wgll, const int* lx, const int* nelv);
/* Adjoint forcing (before the adjoint advection), per component c:
 *   f_c = (obj_scale cc_c + (K chi) ub_c) B
 * with cc = curl(curl ub) mass-averaged (adjoint_minimum_dissipation_source_term.f90:229-243),
 * K chi ub the adjoint lube term (adjoint_lube_source_term.f90:189-203), and the
 * opcolv by B of adjoint_pnpn.f90:671-675. */
int nekoamd_adjoint_forcing(void* fx, void* fy, void* fz, const void* cx,
                            const void* cy, const void* cz, const void* ub,
                            const void* vb, const void* wb, const void* chi,
                            const void* B, const double* K,
                            const double* obj_scale, const int64_t* n);
/* Adjoint advection about the base flow (ub, vb, wb), the device branch of
 * adjoint_advection_no_dealias (adv_adjoint_no_dealias.f90:119-255, :269-303):
 *   f_c -= sum_d ua_d opgrad_c(ub_d);  f_c -= cdtp(ua_c ub, ua_c vb, ua_c wb).
 * (The reference's CPU branch adds the y and z cdtp terms instead, :344; the
 * device branch is the discrete adjoint of the linearised advection.)
 * work: 12 n device doubles. */
int nekoamd_adjoint_advection(void* fx, void* fy, void* fz, const void* ua,
                              const void* va, const void* wa, const void* ub,
                              const void* vb, const void* wb, const void* x,
                              const void* y, const void* z, const void* dx,
                              const void* wgll, const int* lx,
                              const int* nelv, void* work);
/* ---- dealiased advection (over-integration on Gauss-Legendre points) ----
 * The state of adv_lin_dealias_t (sources/adjoint/adv_adjoint_dealias.f90):
 * init_dealias (:138-213) builds the GL space of lxd points per direction,
 * the GLL->GL interpolator and coef_GL with the GLL cofactors interpolated to
 * GL (:155-163); create does the same from the GLL coordinates.  lxd is the
 * reference's default 3 lx / 2 (advection_adjoint_fctry.f90:66-70); lx must
 * be 4, 6 or 8 (kErrUnsupported otherwise). */
int nekoamd_dealias_create(void** st, const void* x, const void* y,
                           const void* z, const void* dx, const int* lx,
                           const int* lxd, const int* nelv);
int nekoamd_dealias_free(void* st);
/* copy the GL metric, [9][nelv lxd^3] with row 3b + a = c_b[a], into cof */
int nekoamd_dealias_metric(const void* st, void* cof);
/* Forward convective term, Neko adv_dealias_t%compute (Neko, not vendored):
 *   f_c = beta f_c - J^T [ t . opgrad(t_c) ],  t = J u on the GL points.
 * beta = 0 overwrites f (f is not read), beta = 1 subtracts as Neko does. */
int nekoamd_dealias_advection(void* st, void* fx, void* fy, void* fz,
                              const void* u, const void* v, const void* w,
                              const double* beta);
/* compute_linear_advection_dealias (adv_adjoint_dealias.f90:466-653):
 *   f_c -= J^T [ u' . opgrad(U_c) ] + J^T [ U . opgrad(u'_c) ] */
int nekoamd_dealias_linear(void* st, void* fx, void* fy, void* fz,
                           const void* up, const void* vp, const void* wp,
                           const void* ub, const void* vb, const void* wb);
/* compute_adjoint_advection_dealias (adv_adjoint_dealias.f90:227-449):
 *   f_c -= J^T [ sum_d ua_d opgrad_c(U_d) ]
 *   f_c -= J^T [ cdtp_x(ua_c U) + cdtp_y(ua_c V) + cdtp_z(ua_c W) ] */
int nekoamd_dealias_adjoint(void* st, void* fx, void* fy, void* fz,
                            const void* ua, const void* va, const void* wa,
                            const void* ub, const void* vb, const void* wb);
/* int_face q u.n on the faces flagged per element in face_bits (bit f =
 * 2*dir + hi); the Dirichlet-velocity surface term of pnpn_prs_res. */
int nekoamd_surface_flux(void* out, const void* u, const void* v,
                         const void* w, const void* x, const void* y,
                         const void* z, const void* dx, const void* wgll,
                         const void* face_bits, const int* lx,
                         const int* nelv);
int nekoamd_box_face_bits(const int* nx, const int* ny, const int* nz,
                          const int* faces, int* bits);
/* sumab%compute_fluid (adjoint_pnpn.f90:665-666): ue = ext[0] u + ext[1] ulag */
int nekoamd_pnpn_extrapolate(void* ue, void* ve, void* we, const void* u,
                             const void* v, const void* w, const void* ul,
                             const void* vl, const void* wl,
                             const double* ext, const int64_t* n);
/* makeabf + makebdf + lag update for one component (adjoint_pnpn.f90:688-700):
 * F = c0 f + c1 flag + B (c2 u + c3 ulag); flag = f; ulag = u */
int nekoamd_pnpn_makef(void* F, void* flag, void* ulag, const void* f,
                       const void* u, const void* B, const double* coef,
                       const int64_t* n);
/* the three components of nekoamd_pnpn_makef in one pass (B read once);
 * each component's results are the one-component call's bit for bit */
int nekoamd_pnpn_makef_vector(void* F1, void* F2, void* F3, void* flag1,
                              void* flag2, void* flag3, void* ulag1,
                              void* ulag2, void* ulag3, const void* f1,
                              const void* f2, const void* f3, const void* u1,
                              const void* u2, const void* u3, const void* B,
                              const double* coef, const int64_t* n);
/* time_order 3 (the shipped cases' case.numerics.time_order, Neko ext_bdf
 * at order 3): the extrapolation ue = ext1 u + ext2 ulag + ext3 ulag2, and
 * makeabf + makebdf with the second lags, F = ext1 f + ext2 flag +
 * ext3 flag2 + (B rho/dt)(bd1 u + bd2 ulag + bd3 ulag2), then flag2 = flag,
 * flag = f, ulag2 = ulag, ulag = u (adjoint_pnpn.f90:665-666, :688-700);
 * coef = {ext1, ext2, ext3, rho/dt bd1, rho/dt bd2, rho/dt bd3}. */
int nekoamd_pnpn_extrapolate3(void* ue, void* ve, void* we, const void* u,
                              const void* v, const void* w, const void* ul,
                              const void* vl, const void* wl, const void* ul2,
                              const void* vl2, const void* wl2,
                              const double* ext, const int64_t* n);
int nekoamd_pnpn_makef3(void* F, void* flag, void* flag2, void* ulag,
                        void* ulag2, const void* f, const void* u,
                        const void* B, const double* coef, const int64_t* n);
/* pressure forcing t = F/rho - mu B c - chi B ue (one component) */
int nekoamd_pnpn_prs_force(void* t, const void* F, const void* c,
                           const void* ue, const void* chi, const void* B,
                           const double* mu, const double* rho,
                           const int64_t* n);
/* the three components of nekoamd_pnpn_prs_force in one pass (pnpn_prs_res's
 * forcing for x, y, z, sources/adjoint/adjoint_pnpn.f90:716-723); each t_c
 * is the one-component call's result bit for bit */
int nekoamd_pnpn_prs_force_vector(void* t1, void* t2, void* t3, const void* F1,
                                  const void* F2, const void* F3,
                                  const void* c1, const void* c2,
                                  const void* c3, const void* ue1,
                                  const void* ue2, const void* ue3,
                                  const void* chi, const void* B,
                                  const double* mu, const double* rho,
                                  const int64_t* n);
/* a = a - b - s c  (c may be NULL: a = a - b) */
int nekoamd_sub2s(void* a, const void* b, const void* c, const double* s,
                  const int64_t* n);
/* r = F - g - r */
int nekoamd_pnpn_vel_res(void* r, const void* F, const void* g,
                         const int64_t* n);
/* U[0,1) per global point of the box lattice (SplitMix64(seed, gid)):
 * the seeded design of the workloads. */
int nekoamd_box_rng_uniform(void* out, const int* nx, const int* ny,
                            const int* nz, const int* lx,
                            const uint64_t* seed);

/* ---- field output: Neko's fld format ------------------------------------
 * One file of a fld series, the file Neko's fld_file_output_t%sample writes
 * for the reference's outputs (steady_state_problem.f90:228-244,
 * topopt_design.f90:283-286): 132-byte header, float32 6.54321, int32 element
 * numbers, element blocks in wdsize (4 or 8) byte words, float32 per-element
 * min/max metadata.  nfields fields map as Neko's writer does: 1 P, 2 P T,
 * 3 U, 4 P U, 5 P U T, n > 5 P U T S(n-5); x, y, z (or null) add the mesh.
 * fields: nfields pointers, device memory when device = 1 (host otherwise;
 * host writes need no GPU). */
int nekoamd_fld_write(const char* path, const double* time, const int* step,
                      const int* lx, const int* nelv, const int* nfields,
                      const void* const* fields, const void* x, const void* y,
                      const void* z, const int* wdsize, const int* device);
/* The series index <name>.nek5000 (filetemplate, firsttimestep,
 * numtimesteps), as examples/power_iterations/baseflow0.nek5000. */
int nekoamd_fld_series(const char* path, const char* prefix, const int* first,
                       const int* count);

/* ---- checkpoint files -----------------------------------------------------
 * The state a Pn/Pn run restarts from, as the reference's chkp_t keeps it
 * (sources/adjoint/adjoint_scheme.f90:159; read back by adjoint_pnpn_restart,
 * sources/adjoint/adjoint_pnpn.f90:418-546).  The layout is this library's
 * (Neko's chkp format is Neko's and no file of it is in the reference):
 * "NKAMDCHK", int32 version 1, int32 nfields, int64 n, int32 tstep, int32 0,
 * float64 time, dt, then nfields x n float64.  fields: nfields device
 * pointers of n values each. */
int nekoamd_chkp_write(const char* path, const int* tstep, const double* time,
                       const double* dt, const int* nfields,
                       const void* const* fields, const int64_t* n);
/* Reads a checkpoint into nfields device fields; nfields and n must match
 * the file's. */
int nekoamd_chkp_read(const char* path, int* tstep, double* time, double* dt,
                      const int* nfields, void* const* fields,
                      const int64_t* n);

#ifdef __cplusplus
}
#endif
#endif
