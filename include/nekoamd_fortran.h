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

#ifdef __cplusplus
}
#endif
#endif
