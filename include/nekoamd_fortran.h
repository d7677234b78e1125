/*
 * nekoamd_fortran.h -- bind(c) entry points of the Fortran host library
 * libnekoamd_f.so (neko-top_amd/fortran/).  The Fortran host implements the
 * reference's solver plugin surface (ax_t / pc_t / ksp_t / source_term_t,
 * factories and type-bound procedures) over the C-ABI of nekoamd.h; these
 * exports let a C or Python caller drive whole solves through it.
 * Scalars by reference, arrays as host pointers, status 0 = success.
 */
#ifndef NEKOAMD_FORTRAN_H
#define NEKOAMD_FORTRAN_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Jacobi-PCG through ax_helm_factory / precon_factory('jacobi') /
 * krylov_solver_factory('cg') / set_pc / solve on an nx*ny*nz box of order
 * lx-1 (replaces ksp_prs%solve, sources/adjoint/adjoint_pnpn.f90:734-736,
 * kind 0, and one component of ksp_vel%solve_coupled, :784-786, kind 1).
 * f and x hold nx*ny*nz*lx^3 values; f is gs-assembled and masked inside.
 * res[0] = start residual, res[1] = final residual. */
int nekoamd_f_pcg_solve(const int* nx, const int* ny, const int* nz,
                        const int* lx, const double* jitter,
                        const int64_t* seed, const int* kind, const double* re,
                        const double* dt, const double* f, double* x,
                        const double* abstol, const int* max_iter, int* iters,
                        double* res);

/* Brinkman channel (fluid_pnpn_t of the Fortran host): one Pn/Pn step in the
 * order of sources/adjoint/adjoint_pnpn.f90:619-827 per call of _step.
 * create: nx*ny*nz box of order lx-1, uniform inflow at x = 0, outflow at
 *   x = L, no-slip walls; Jacobi-PCG tolerances abstol_p / abstol_v.
 * design: rho ~ U[0,1) per global point from seed, npass PDE-filter passes
 *   (PDE_filter_mapping.f90:212-278), chi = RAMP convex-up(f_min, f_max, q)
 *   (RAMP_mapping.f90:227-241) -- the topopt_design_t%map_forward chain.
 * step: iters[4*s..4*s+3], res[...] = pressure, u, v, w PCG iterations and
 *   final residuals of step s.
 * get / set: which = 0 u, 1 v, 2 w, 3 p, 4 chi, 5 rho, 6 filtered rho,
 *   7-9 the base flow ub, vb, wb of an adjoint case, 10-12 the velocity
 *   lags, 13-15 the lagged explicit forcing, 16-21 their second lags
 *   (time_order 3). */
int nekoamd_f_channel_create(const int* nx, const int* ny, const int* nz,
                             const int* lx, const double* jitter,
                             const int64_t* seed, const double* re,
                             const double* dt, const double* abstol_p,
                             const double* abstol_v, const int* max_iter,
                             const int* precision, int* handle);
/* Projection spaces (the reference's proj_prs / proj_u,v,w): L_prs and
 * L_vel previous solutions (0 = off, at most 64), active after activ_step
 * steps; keep_prs / keep_vel solutions kept when a full basis restarts
 * (0: Neko's restart from the newest solution; otherwise below L).
 * Returns 1 on a bad size. */
int nekoamd_f_channel_projection(const int* handle, const int* L_prs,
                                 const int* L_vel, const int* activ_step,
                                 const int* keep_prs, const int* keep_vel);
/* The channel's projection bases in FP32 (*on != 0) or FP64, after
 * nekoamd_f_channel_projection and before the first step
 * (nekoamd_proj_set_fp32 on each). */
int nekoamd_f_channel_projection_fp32(const int* handle, const int* on);
/* Learned guess of the pressure solves (nekoamd_mlp_*): mode 0 off,
 * 1 harvest pairs, 2 apply, 3 apply and harvest.  The first call creates the
 * network (capacity pairs, batch training rows, weight seed). */
int nekoamd_f_channel_learned(const int* handle, const int* mode,
                              const int* capacity, const int* batch,
                              const int64_t* seed);
/* The network's correction inside a flexible-CG pressure solve every
 * `every`-th iteration (0: off): fluid_pnpn_t passes the channel's network to
 * the 'fcg' pressure solver (nekoamd_f_channel_pressure_solver(h, "fcg",
 * "jacobi")).  No reference counterpart. */
int nekoamd_f_channel_learned_correction(const int* handle, const int* every);
/* Train on the harvested pairs; mean loss of the first / last epoch. */
int nekoamd_f_channel_train(const int* handle, const int* epochs,
                            const int* batch, const float* lr,
                            const int64_t* seed, double* loss0, double* loss1,
                            int* npairs);
/* gamma of the last pressure solve's learned direction (0: none). */
int nekoamd_f_channel_gamma(const int* handle, double* gamma);
/* Chain rule of the design maps (topopt_design.f90:316-352): sens =
 * filter^T(RAMP'(rho~) df_dchi), host arrays of n values.  Needs a design. */
int nekoamd_f_channel_map_backward(const int* handle, const double* df_dchi,
                                   double* sens);
/* Turn the channel into the adjoint solve of adjoint_pnpn_step
 * (adjoint_pnpn.f90:619-827): adjoint state zeroed, forcing
 * obj_scale curl(curl ub) + K chi ub, adjoint advection about the base flow
 * (set it as fields 7-9). */
int nekoamd_f_channel_adjoint(const int* handle, const double* K,
                              const double* obj_scale);
/* The Brinkman source term (simple_brinkman_source_term.f90:139-153) of the
 * channel's step: implicit = 1 (default) semi-implicit, chi in the velocity
 * Helmholtz operator and chi B u_e in the pressure forcing; implicit = 0 the
 * reference's explicit treatment f -= chi u with EXT
 * (adjoint_pnpn.f90:669-691), stable for chi dt < 4/3. */
int nekoamd_f_channel_brinkman(const int* handle, const int* implicit);
/* case.numerics.dealias (advection_adjoint_fctry.f90:57-99): on = 1 computes
 * the forward and adjoint advection terms on lxd Gauss-Legendre points per
 * direction (lxd <= 0: 3 (p + 1) / 2), on = 0 with the collocated forms. */
int nekoamd_f_channel_dealias(const int* handle, const int* on,
                              const int* lxd);
/* The optimisation loop of examples/rugby_ball/driver.f90:69-138 (with
 * steady_state_problem.f90:273-353) on a channel case: forward solve,
 * objective and volume constraint, adjoint solve, sensitivity, chain rule,
 * MMA update (constraint scaled by 100), KKT, map_forward.  precision:
 * NK_CG_FP64 / NK_CG_MIXED solves; the design indicator starts as seeded
 * noise smoothed by npass-1 filter passes. */
int nekoamd_f_topopt_create(const int* nx, const int* ny, const int* nz,
                            const int* lx, const double* jitter,
                            const int64_t* seed, const double* re,
                            const double* dt, const double* abstol_p,
                            const double* abstol_v, const int* max_iter,
                            const int* precision, const int64_t* design_seed,
                            const int* npass, const double* f_min,
                            const double* f_max, const double* q,
                            const double* v_max, const int* is_max,
                            int* handle);
/* One design iteration: out[0..7] = F, dissipation, lube, volume constraint,
 * volume fraction, KKT residumax, residunorm, MMA Newton iterations;
 * iters[0..1] = pressure iterations of the forward / adjoint solves. */
int nekoamd_f_topopt_iterate(const int* handle, const int* nfwd,
                             const int* nadj, double* out, int* iters);
/* which: 0 design indicator, 1 filtered design, 2 chi, 3 dF/drho,
 * 4 dF/dchi, 5-7 forward u, v, w, 8-10 adjoint u, v, w (n host values). */
int nekoamd_f_topopt_get(const int* handle, const int* which, double* out);
int nekoamd_f_topopt_free(const int* handle);
/* fld output (Neko's fld_file_output_t): sample p, u, v, w of the channel
 * into the series `name` (name0.f00000, ..., name.nek5000) at time t; a new
 * name or wdsize (4 or 8) starts a new series. */
int nekoamd_f_channel_sample(const int* handle, const char* name,
                             const double* t, const int* wdsize);
/* the loop's 13-field 'optimization' series (steady_state_problem.f90:
 * 216-244), sampled every design iteration after the sensitivities */
int nekoamd_f_topopt_output(const int* handle, const char* name,
                            const int* wdsize);
/* The pressure solver of the channel (case.fluid.pressure_solver): solver
 * "cg" | "cg_mixed" | "gmres", pc "jacobi" | "hsmg"; gmres + hsmg is what
 * the shipped cases run (rugby_ball.case:46-52).  Tolerance and iteration cap
 * are kept. */
int nekoamd_f_channel_pressure_solver(const int* handle, const char* solver,
                                      const char* pc);
int nekoamd_f_topopt_pressure_solver(const int* handle, const char* solver,
                                     const char* pc);
/* nekoamd_f_channel_dealias for both solves of the loop */
int nekoamd_f_topopt_dealias(const int* handle, const int* on, const int* lxd);
int nekoamd_f_channel_design(const int* handle, const int64_t* seed,
                             const int* npass, const double* f_min,
                             const double* f_max, const double* q);
int nekoamd_f_channel_step(const int* handle, const int* nsteps, int* iters,
                           double* res);
int nekoamd_f_channel_get(const int* handle, const int* which, double* out);
int nekoamd_f_channel_set(const int* handle, const int* which,
                          const double* in);
/* case.numerics.time_order of a channel that has not stepped yet: 2
 * (BDF2/EXT2 after a BDF1/EXT1 first step) or 3 (BDF3/EXT3 from the third
 * step, Neko's ext_bdf at order 3 as the reference's cases run). */
int nekoamd_f_channel_time_order(const int* handle, const int* order);
/* Step counter of the channel: set = 0 reads it into *tstep, set = 1 sets
 * it (the next step is tstep + 1). */
int nekoamd_f_channel_tstep(const int* handle, int* tstep, const int* set);
/* fluid_pnpn_t%checkpoint (restart = 0) or %restart (restart = 1): the
 * state u, v, w, p, the velocity lags and the lagged forcing, and the step
 * counter, in a nekoamd_chkp_write file at path (NUL-terminated); the
 * restart makes u, v, w, p and the lags continuous again as
 * adjoint_pnpn_restart does (sources/adjoint/adjoint_pnpn.f90:418-484). */
int nekoamd_f_channel_chkp(const int* handle, const char* path,
                           const int* restart);
int nekoamd_f_channel_free(const int* handle);

#ifdef __cplusplus
}
#endif
#endif
