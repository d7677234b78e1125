"""ctypes signatures of oracle/topopt.c and oracle/mma.c (test infrastructure)."""
import ctypes as C


def declare(L, dp, i32p):
    i64, dbl, I = C.c_int64, C.c_double, C.c_int
    vp = C.c_void_p
    L.or_brinkman.argtypes = [i64, dp, dp, dp, dp, dp, dp, dp]
    L.or_update_permeability.argtypes = [i64, dp, dp, dbl, dbl, dbl]
    L.or_col3_mask.argtypes = [dp, dp, dp, vp, I]
    L.or_ramp_fwd.argtypes = [i64, dp, dp, dbl, dbl, dbl, I]
    L.or_ramp_bwd.argtypes = [i64, dp, dp, dp, dbl, dbl, dbl, I]
    L.or_linear_fwd.argtypes = [i64, dp, dp, dbl, dbl]
    L.or_linear_bwd.argtypes = [i64, dp, dp, dbl, dbl]
    L.or_grad.argtypes = [I, I, dp, dp, dp, dp, dp, dp, dp, dp]
    L.or_dissipation_objective.argtypes = [I, I, dp, dp, dp, dp, dp, dp, dp, dp, dp,
                                           dbl, dbl, I, dp]
    L.or_dissipation_sensitivity.argtypes = [i64, dp, dp, dp, dp, dp, dp, dp, dbl, dbl, I]
    L.or_adjoint_lube.argtypes = [i64, dp, dp, dp, dp, dp, dp, dp, dbl]
    L.or_volume_constraint.argtypes = [i64, dp, dp, dbl, dbl, I, dp]
    L.or_volume_sensitivity.argtypes = [i64, dp, dbl, I]
    L.or_filter_apply.restype = I
    L.or_filter_apply.argtypes = [vp, dbl, dbl, I, dp, dp]
    L.or_mma_create.restype = vp
    L.or_mma_create.argtypes = [I, I, dp, dbl, dp, dp, dp, dp, dp, I, dbl, dbl, dbl, dbl]
    L.or_mma_free.argtypes = [vp]
    L.or_mma_update.restype = I
    L.or_mma_update.argtypes = [vp, I, dp, dp, dp, dp]
    L.or_mma_kkt.argtypes = [vp, dp, dp, dp, dp]
    L.or_mma_residumax.restype = dbl
    L.or_mma_residumax.argtypes = [vp]
    L.or_mma_residunorm.restype = dbl
    L.or_mma_residunorm.argtypes = [vp]
    L.or_mma_get_subproblem.argtypes = [vp] + [dp] * 9
    L.or_mma_get_duals.argtypes = [vp] + [dp] * 7
    L.or_dgesv.restype = I
    L.or_dgesv.argtypes = [I, dp, dp]


def declare_pnpn(L, dp):
    I = C.c_int
    L.or_advection.argtypes = [I, I, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp]
    L.or_curl_local.argtypes = [I, I, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp]
    L.or_opgrad.argtypes = [I, I, dp, dp, dp, dp, dp, dp, dp, dp, dp]
    L.or_cdtp.argtypes = [I, I, dp, dp, dp, dp, dp, dp, dp, dp, dp]
    L.or_surface_flux.argtypes = [I, I, I, I, dp, dp, dp, dp, dp, dp, dp, dp, I, dp]
    L.or_pnpn_coeffs.argtypes = [I, dp, dp]
    L.or_pnpn_coeffs3.argtypes = [I, I, dp, dp]
    L.or_pnpn_step.argtypes = [C.c_void_p]
    L.or_adjoint_advection.argtypes = [I, I] + [dp] * 15
    L.or_cofactors.argtypes = [I, I, dp, dp, dp, dp, dp]
    L.or_gauss_legendre.argtypes = [I, dp, dp]
    L.or_lagrange_matrix.argtypes = [I, dp, I, dp, dp]
    L.or_deriv_matrix.argtypes = [I, dp, dp]
    L.or_dealias_metric.argtypes = [I, I, I, dp, dp, dp]
    L.or_dealias_advection.argtypes = [I, I, I] + [dp] * 10
    L.or_dealias_linear.argtypes = [I, I, I] + [dp] * 13
    L.or_dealias_adjoint.argtypes = [I, I, I] + [dp] * 13
