!> Fortran interfaces to the C-ABI of libnekoamd (include/nekoamd.h).
!!
!! Same seam as the reference's own device wrappers
!! (sources/neko_ext/math/bcknd/device_math_ext.f90:39-105): device buffers
!! are type(c_ptr), value; scalars are passed by reference; every entry
!! returns an integer status that nk_check turns into an abort with the
!! library's message (the reference aborts via CUDA_CHECK, math_ext.cu:56).
module nekoamd_capi
  use, intrinsic :: iso_c_binding
  implicit none
  public

  integer, parameter :: rp = c_double
  integer(c_int), parameter :: NK_HOST_TO_DEVICE = 1, NK_DEVICE_TO_HOST = 2, &
       NK_DEVICE_TO_DEVICE = 3
  integer(c_int), parameter :: GS_OP_ADD = 1, GS_OP_MUL = 2, GS_OP_MIN = 3, &
       GS_OP_MAX = 4
  !> nekoamd_cg_create precisions (NEKOAMD_CG_FP64 / NEKOAMD_CG_MIXED)
  integer(c_int), parameter :: NK_CG_FP64 = 0, NK_CG_MIXED = 1
  integer(c_int), parameter :: NK_FACE_XMIN = 1, NK_FACE_XMAX = 2, &
       NK_FACE_YMIN = 4, NK_FACE_YMAX = 8, NK_FACE_ZMIN = 16, NK_FACE_ZMAX = 32

  type, bind(c) :: nk_cg_monitor_c
     integer(c_int) :: iter
     integer(c_int) :: done
     real(c_double) :: res_start
     real(c_double) :: res_final
     integer(c_int) :: nupdate
     integer(c_int) :: pad
  end type nk_cg_monitor_c

  interface
     function nekoamd_error_string() result(p) bind(c, name='nekoamd_error_string')
       import c_ptr
       type(c_ptr) :: p
     end function

     integer(c_int) function nekoamd_device_init(dev) bind(c, name='nekoamd_device_init')
       import c_int
       integer(c_int) :: dev
     end function

     integer(c_int) function nekoamd_device_alloc(p, bytes) bind(c, name='nekoamd_device_alloc')
       import c_int, c_ptr, c_int64_t
       type(c_ptr) :: p
       integer(c_int64_t) :: bytes
     end function

     integer(c_int) function nekoamd_device_free(p) bind(c, name='nekoamd_device_free')
       import c_int, c_ptr
       type(c_ptr), value :: p
     end function

     integer(c_int) function nekoamd_device_memcpy(dst, src, bytes, kind, sync) &
          bind(c, name='nekoamd_device_memcpy')
       import c_int, c_ptr, c_int64_t
       type(c_ptr), value :: dst, src
       integer(c_int64_t) :: bytes
       integer(c_int) :: kind, sync
     end function

     integer(c_int) function nekoamd_device_sync() bind(c, name='nekoamd_device_sync')
       import c_int
     end function

     integer(c_int) function nekoamd_gll(lx, z, w, dx) bind(c, name='nekoamd_gll')
       import c_int, c_double
       integer(c_int) :: lx
       real(c_double) :: z(*), w(*), dx(*)
     end function

     integer(c_int) function nekoamd_box_vertices(nx, ny, nz, lxyz, jitter, seed, v) &
          bind(c, name='nekoamd_box_vertices')
       import c_int, c_double, c_int64_t
       integer(c_int) :: nx, ny, nz
       real(c_double) :: lxyz(3), jitter
       integer(c_int64_t) :: seed
       real(c_double) :: v(*)
     end function

     integer(c_int) function nekoamd_box_corners(nx, ny, nz, v, corners) &
          bind(c, name='nekoamd_box_corners')
       import c_int, c_double
       integer(c_int) :: nx, ny, nz
       real(c_double) :: v(*), corners(*)
     end function

     integer(c_int) function nekoamd_box_gids(nx, ny, nz, lx, gid) &
          bind(c, name='nekoamd_box_gids')
       import c_int, c_int64_t
       integer(c_int) :: nx, ny, nz, lx
       integer(c_int64_t) :: gid(*)
     end function

     integer(c_int) function nekoamd_box_face_mask(nx, ny, nz, lx, faces, out, cnt) &
          bind(c, name='nekoamd_box_face_mask')
       import c_int, c_int64_t, c_ptr
       integer(c_int) :: nx, ny, nz, lx, faces
       type(c_ptr), value :: out
       integer(c_int64_t) :: cnt
     end function

     integer(c_int) function nekoamd_coef_geo(x, y, z, g11, g12, g13, g22, g23, g33, &
          b, jac, corners, zg, wg, dx, lx, nelv) bind(c, name='nekoamd_coef_geo')
       import c_int, c_ptr
       type(c_ptr), value :: x, y, z, g11, g12, g13, g22, g23, g33, b, jac, &
            corners, zg, wg, dx
       integer(c_int) :: lx, nelv
     end function

     integer(c_int) function nekoamd_ax_helm_f64(w, u, dx, h1, h2, g11, g12, g13, g22, &
          g23, g33, b, lx, nelv, ifh2) bind(c, name='nekoamd_ax_helm_f64')
       import c_int, c_ptr
       type(c_ptr), value :: w, u, dx, h1, h2, g11, g12, g13, g22, g23, g33, b
       integer(c_int) :: lx, nelv, ifh2
     end function

     integer(c_int) function nekoamd_ax_helm_vector_f64(w1, w2, w3, u1, u2, u3, dx, h1, &
          h2, g11, g12, g13, g22, g23, g33, b, lx, nelv, ifh2) &
          bind(c, name='nekoamd_ax_helm_vector_f64')
       import c_int, c_ptr
       type(c_ptr), value :: w1, w2, w3, u1, u2, u3, dx, h1, h2, g11, g12, g13, g22, &
            g23, g33, b
       integer(c_int) :: lx, nelv, ifh2
     end function

     integer(c_int) function nekoamd_pnpn_vel_res_vector(r1, r2, r3, u1, u2, u3, f1, f2, &
          f3, g1, g2, g3, dx, h1, h2, g11, g12, g13, g22, g23, g33, b, lx, nelv, ifh2) &
          bind(c, name='nekoamd_pnpn_vel_res_vector')
       import c_int, c_ptr
       type(c_ptr), value :: r1, r2, r3, u1, u2, u3, f1, f2, f3, g1, g2, g3, dx, h1, h2, &
            g11, g12, g13, g22, g23, g33, b
       integer(c_int) :: lx, nelv, ifh2
     end function

     integer(c_int) function nekoamd_gs_init(gs, gid, n) bind(c, name='nekoamd_gs_init')
       import c_int, c_ptr, c_int64_t
       type(c_ptr) :: gs
       integer(c_int64_t) :: gid(*)
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_gs_free(gs) bind(c, name='nekoamd_gs_free')
       import c_int, c_ptr
       type(c_ptr), value :: gs
     end function

     integer(c_int) function nekoamd_gs_op_f64(gs, u, op) bind(c, name='nekoamd_gs_op_f64')
       import c_int, c_ptr
       type(c_ptr), value :: gs, u
       integer(c_int) :: op
     end function

     integer(c_int) function nekoamd_gs_mass_average(gs, u, b, binv) &
          bind(c, name='nekoamd_gs_mass_average')
       import c_int, c_ptr
       type(c_ptr), value :: gs, u, b, binv
     end function

     integer(c_int) function nekoamd_copy(a, b, n) bind(c, name='nekoamd_copy')
       import c_int, c_ptr, c_int64_t
       type(c_ptr), value :: a, b
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_rzero(a, n) bind(c, name='nekoamd_rzero')
       import c_int, c_ptr, c_int64_t
       type(c_ptr), value :: a
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_cfill(a, c, n) bind(c, name='nekoamd_cfill')
       import c_int, c_ptr, c_int64_t, c_double
       type(c_ptr), value :: a
       real(c_double) :: c
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_cmult(a, c, n) bind(c, name='nekoamd_cmult')
       import c_int, c_ptr, c_int64_t, c_double
       type(c_ptr), value :: a
       real(c_double) :: c
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_add2(a, b, n) bind(c, name='nekoamd_add2')
       import c_int, c_ptr, c_int64_t
       type(c_ptr), value :: a, b
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_cadd2(a, b, c, n) bind(c, name='nekoamd_cadd2')
       import c_int, c_ptr, c_int64_t, c_double
       type(c_ptr), value :: a, b
       real(c_double) :: c
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_add2s2(a, b, c, n) bind(c, name='nekoamd_add2s2')
       import c_int, c_ptr, c_int64_t, c_double
       type(c_ptr), value :: a, b
       real(c_double) :: c
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_col2(a, b, n) bind(c, name='nekoamd_col2')
       import c_int, c_ptr, c_int64_t
       type(c_ptr), value :: a, b
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_col3(a, b, c, n) bind(c, name='nekoamd_col3')
       import c_int, c_ptr, c_int64_t
       type(c_ptr), value :: a, b, c
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_addcol3(a, b, c, n) bind(c, name='nekoamd_addcol3')
       import c_int, c_ptr, c_int64_t
       type(c_ptr), value :: a, b, c
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_subcol3(a, b, c, n) bind(c, name='nekoamd_subcol3')
       import c_int, c_ptr, c_int64_t
       type(c_ptr), value :: a, b, c
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_invcol1(a, n) bind(c, name='nekoamd_invcol1')
       import c_int, c_ptr, c_int64_t
       type(c_ptr), value :: a
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_bc_zero(u, mask, nmask) bind(c, name='nekoamd_bc_zero')
       import c_int, c_ptr, c_int64_t
       type(c_ptr), value :: u, mask
       integer(c_int64_t) :: nmask
     end function

     integer(c_int) function nekoamd_glsc2(res, a, b, n) bind(c, name='nekoamd_glsc2')
       import c_int, c_ptr, c_int64_t, c_double
       real(c_double) :: res
       type(c_ptr), value :: a, b
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_glsc3(res, a, b, c, n) bind(c, name='nekoamd_glsc3')
       import c_int, c_ptr, c_int64_t, c_double
       real(c_double) :: res
       type(c_ptr), value :: a, b, c
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_jacobi_update(d, gs, dx, g11, g22, g33, h1, h2, b, &
          lx, nelv, ifh2) bind(c, name='nekoamd_jacobi_update')
       import c_int, c_ptr
       type(c_ptr), value :: d, gs, dx, g11, g22, g33, h1, h2, b
       integer(c_int) :: lx, nelv, ifh2
     end function

     integer(c_int) function nekoamd_jacobi_solve(z, r, d, n) bind(c, name='nekoamd_jacobi_solve')
       import c_int, c_ptr, c_int64_t
       type(c_ptr), value :: z, r, d
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_cg_create(cg, n, nelv, precision) &
          bind(c, name='nekoamd_cg_create')
       import c_int, c_ptr, c_int64_t
       type(c_ptr) :: cg
       integer(c_int64_t) :: n
       integer(c_int) :: nelv, precision
     end function

     integer(c_int) function nekoamd_cg_free(cg) bind(c, name='nekoamd_cg_free')
       import c_int, c_ptr
       type(c_ptr), value :: cg
     end function

     integer(c_int) function nekoamd_cg_solve3(cg, x, f, d, mult, volume, abstol, &
          max_iter, poll, gs, mask, nmask, dx, h1, h2, g11, g12, g13, g22, g23, g33, b, &
          lx, ifh2, mon) bind(c, name='nekoamd_cg_solve3')
       import c_int, c_ptr, c_int64_t, c_double, nk_cg_monitor_c
       type(c_ptr) :: cg(3), x(3), f(3)
       type(c_ptr), value :: d, mult
       real(c_double) :: volume, abstol
       integer(c_int) :: max_iter, poll
       type(c_ptr), value :: gs, mask
       integer(c_int64_t) :: nmask
       type(c_ptr), value :: dx, h1, h2, g11, g12, g13, g22, g23, g33, b
       integer(c_int) :: lx, ifh2
       type(nk_cg_monitor_c) :: mon(3)
     end function

     integer(c_int) function nekoamd_cg_solve(cg, x, f, d, mult, volume, abstol, &
          max_iter, poll, gs, mask, nmask, dx, h1, h2, g11, g12, g13, g22, g23, g33, b, &
          lx, ifh2, mon) bind(c, name='nekoamd_cg_solve')
       import c_int, c_ptr, c_int64_t, c_double, nk_cg_monitor_c
       type(c_ptr), value :: cg, x, f, d, mult
       real(c_double) :: volume, abstol
       integer(c_int) :: max_iter, poll
       type(c_ptr), value :: gs, mask
       integer(c_int64_t) :: nmask
       type(c_ptr), value :: dx, h1, h2, g11, g12, g13, g22, g23, g33, b
       integer(c_int) :: lx, ifh2
       type(nk_cg_monitor_c) :: mon
     end function

     integer(c_int) function nekoamd_proj_create(proj, n, l) bind(c, name='nekoamd_proj_create')
       import c_int, c_ptr, c_int64_t
       type(c_ptr) :: proj
       integer(c_int64_t) :: n
       integer(c_int) :: l
     end function

     integer(c_int) function nekoamd_proj_free(proj) bind(c, name='nekoamd_proj_free')
       import c_int, c_ptr
       type(c_ptr), value :: proj
     end function

     integer(c_int) function nekoamd_proj_clear(proj) bind(c, name='nekoamd_proj_clear')
       import c_int, c_ptr
       type(c_ptr), value :: proj
     end function

     integer(c_int) function nekoamd_proj_size(proj, m) bind(c, name='nekoamd_proj_size')
       import c_int, c_ptr
       type(c_ptr), value :: proj
       integer(c_int) :: m
     end function

     integer(c_int) function nekoamd_proj_set_keep(proj, keep) &
          bind(c, name='nekoamd_proj_set_keep')
       import c_int, c_ptr
       type(c_ptr), value :: proj
       integer(c_int) :: keep
     end function

     integer(c_int) function nekoamd_proj_set_fp32(proj, on) bind(c, name='nekoamd_proj_set_fp32')
       import c_int, c_ptr
       type(c_ptr), value :: proj
       integer(c_int) :: on
     end function

     integer(c_int) function nekoamd_proj_pre(proj, b, mult) bind(c, name='nekoamd_proj_pre')
       import c_int, c_ptr
       type(c_ptr), value :: proj, b, mult
     end function

     integer(c_int) function nekoamd_proj_augment(proj, b, q, aq, mult, gamma) &
          bind(c, name='nekoamd_proj_augment')
       import c_int, c_ptr, c_double
       type(c_ptr), value :: proj, b, q, aq, mult
       real(c_double) :: gamma
     end function

     integer(c_int) function nekoamd_proj_post(proj, x, mult, gs, mask, nmask, dx, h1, h2, &
          g11, g12, g13, g22, g23, g33, b, lx, nelv, ifh2) bind(c, name='nekoamd_proj_post')
       import c_int, c_ptr, c_int64_t
       type(c_ptr), value :: proj, x, mult, gs, mask
       integer(c_int64_t) :: nmask
       type(c_ptr), value :: dx, h1, h2, g11, g12, g13, g22, g23, g33, b
       integer(c_int) :: lx, nelv, ifh2
     end function

     integer(c_int) function nekoamd_mlp_create(mlp, nx, ny, nz, lx, capacity, batch, seed) &
          bind(c, name='nekoamd_mlp_create')
       import c_int, c_ptr, c_int64_t
       type(c_ptr) :: mlp
       integer(c_int) :: nx, ny, nz, lx, capacity, batch
       integer(c_int64_t) :: seed
     end function

     integer(c_int) function nekoamd_mlp_free(mlp) bind(c, name='nekoamd_mlp_free')
       import c_int, c_ptr
       type(c_ptr), value :: mlp
     end function

     integer(c_int) function nekoamd_mlp_pairs(mlp, npairs) bind(c, name='nekoamd_mlp_pairs')
       import c_int, c_ptr
       type(c_ptr), value :: mlp
       integer(c_int) :: npairs
     end function

     integer(c_int) function nekoamd_mlp_train(mlp, epochs, batch, lr, seed, loss0, loss1) &
          bind(c, name='nekoamd_mlp_train')
       import c_int, c_ptr, c_int64_t, c_float, c_double
       type(c_ptr), value :: mlp
       integer(c_int) :: epochs, batch
       real(c_float) :: lr
       integer(c_int64_t) :: seed
       real(c_double) :: loss0, loss1
     end function

     integer(c_int) function nekoamd_mlp_pre(mlp, b, mult, gs, mask, nmask, dx, h1, h2, g11, &
          g12, g13, g22, g23, g33, bm, lx, ifh2, apply, save, gamma) &
          bind(c, name='nekoamd_mlp_pre')
       import c_int, c_ptr, c_int64_t, c_double
       type(c_ptr), value :: mlp, b, mult, gs, mask
       integer(c_int64_t) :: nmask
       type(c_ptr), value :: dx, h1, h2, g11, g12, g13, g22, g23, g33, bm
       integer(c_int) :: lx, ifh2, apply, save
       real(c_double) :: gamma
     end function

     integer(c_int) function nekoamd_mlp_post(mlp, x, harvest) bind(c, name='nekoamd_mlp_post')
       import c_int, c_ptr
       type(c_ptr), value :: mlp, x
       integer(c_int) :: harvest
     end function

     integer(c_int) function nekoamd_adjoint_forcing(fx, fy, fz, cx, cy, cz, ub, vb, wb, chi, &
          bm, k, obj_scale, n) bind(c, name='nekoamd_adjoint_forcing')
       import c_int, c_ptr, c_int64_t, c_double
       type(c_ptr), value :: fx, fy, fz, cx, cy, cz, ub, vb, wb, chi, bm
       real(c_double) :: k, obj_scale
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_adjoint_advection(fx, fy, fz, ua, va, wa, ub, vb, wb, &
          x, y, z, dx, wgll, lx, nelv, work) bind(c, name='nekoamd_adjoint_advection')
       import c_int, c_ptr
       type(c_ptr), value :: fx, fy, fz, ua, va, wa, ub, vb, wb, x, y, z, dx, wgll, work
       integer(c_int) :: lx, nelv
     end function

     integer(c_int) function nekoamd_gmres_create(st, n, m) bind(c, name='nekoamd_gmres_create')
       import c_int, c_ptr, c_int64_t
       type(c_ptr) :: st
       integer(c_int64_t) :: n
       integer(c_int) :: m
     end function

     integer(c_int) function nekoamd_gmres_free(st) bind(c, name='nekoamd_gmres_free')
       import c_int, c_ptr
       type(c_ptr), value :: st
     end function

     integer(c_int) function nekoamd_gmres_solve(st, x, f, pc_kind, pc, mult, volume, abstol, &
          max_iter, gs, mask, nmask, dx, h1, h2, g11, g12, g13, g22, g23, g33, b, lx, ifh2, &
          mon) bind(c, name='nekoamd_gmres_solve')
       import c_int, c_ptr, c_int64_t, c_double, nk_cg_monitor_c
       type(c_ptr), value :: st, x, f, pc, mult, gs, mask, dx, h1, h2, g11, g12, g13, g22, &
            g23, g33, b
       integer(c_int) :: pc_kind, max_iter, lx, ifh2
       real(c_double) :: volume, abstol
       integer(c_int64_t) :: nmask
       type(nk_cg_monitor_c) :: mon
     end function

     integer(c_int) function nekoamd_fcg_create(st, n, nelv) bind(c, name='nekoamd_fcg_create')
       import c_int, c_ptr, c_int64_t
       type(c_ptr) :: st
       integer(c_int64_t) :: n
       integer(c_int) :: nelv
     end function

     integer(c_int) function nekoamd_fcg_free(st) bind(c, name='nekoamd_fcg_free')
       import c_int, c_ptr
       type(c_ptr), value :: st
     end function

     integer(c_int) function nekoamd_fcg_solve(st, x, f, d, mult, volume, abstol, max_iter, &
          poll, gs, mask, nmask, dx, h1, h2, g11, g12, g13, g22, g23, g33, b, lx, ifh2, &
          mlp, every, mon) bind(c, name='nekoamd_fcg_solve')
       import c_int, c_ptr, c_int64_t, c_double, nk_cg_monitor_c
       type(c_ptr), value :: st, x, f, d, mult, gs, mask, dx, h1, h2, g11, g12, g13, g22, &
            g23, g33, b, mlp
       integer(c_int) :: max_iter, poll, lx, ifh2, every
       real(c_double) :: volume, abstol
       integer(c_int64_t) :: nmask
       type(nk_cg_monitor_c) :: mon
     end function

     integer(c_int) function nekoamd_hsmg_create(st, nx, ny, nz, lx, corners, faces, crs_iter, &
          precision) bind(c, name='nekoamd_hsmg_create')
       import c_int, c_ptr, c_double
       type(c_ptr) :: st
       integer(c_int) :: nx, ny, nz, lx, faces, crs_iter, precision
       real(c_double) :: corners(*)
     end function

     integer(c_int) function nekoamd_hsmg_free(st) bind(c, name='nekoamd_hsmg_free')
       import c_int, c_ptr
       type(c_ptr), value :: st
     end function

     integer(c_int) function nekoamd_hsmg_apply(st, z, r) bind(c, name='nekoamd_hsmg_apply')
       import c_int, c_ptr
       type(c_ptr), value :: st, z, r
     end function

     integer(c_int) function nekoamd_dealias_create(st, x, y, z, dx, lx, lxd, nelv) &
          bind(c, name='nekoamd_dealias_create')
       import c_int, c_ptr
       type(c_ptr) :: st
       type(c_ptr), value :: x, y, z, dx
       integer(c_int) :: lx, lxd, nelv
     end function

     integer(c_int) function nekoamd_dealias_free(st) bind(c, name='nekoamd_dealias_free')
       import c_int, c_ptr
       type(c_ptr), value :: st
     end function

     integer(c_int) function nekoamd_dealias_advection(st, fx, fy, fz, u, v, w, beta) &
          bind(c, name='nekoamd_dealias_advection')
       import c_int, c_ptr, c_double
       type(c_ptr), value :: st, fx, fy, fz, u, v, w
       real(c_double) :: beta
     end function

     integer(c_int) function nekoamd_dealias_linear(st, fx, fy, fz, up, vp, wp, ub, vb, wb) &
          bind(c, name='nekoamd_dealias_linear')
       import c_int, c_ptr
       type(c_ptr), value :: st, fx, fy, fz, up, vp, wp, ub, vb, wb
     end function

     integer(c_int) function nekoamd_dealias_adjoint(st, fx, fy, fz, ua, va, wa, ub, vb, wb) &
          bind(c, name='nekoamd_dealias_adjoint')
       import c_int, c_ptr
       type(c_ptr), value :: st, fx, fy, fz, ua, va, wa, ub, vb, wb
     end function

     integer(c_int) function nekoamd_fld_write(path, time, step, lx, nelv, nfields, fields, &
          x, y, z, wdsize, device) bind(c, name='nekoamd_fld_write')
       import c_int, c_ptr, c_double
       type(c_ptr), value :: path, fields, x, y, z
       real(c_double) :: time
       integer(c_int) :: step, lx, nelv, nfields, wdsize, device
     end function

     integer(c_int) function nekoamd_fld_series(path, prefix, first, count) &
          bind(c, name='nekoamd_fld_series')
       import c_int, c_ptr
       type(c_ptr), value :: path, prefix
       integer(c_int) :: first, count
     end function

     integer(c_int) function nekoamd_chkp_write(path, tstep, time, dt, nfields, fields, n) &
          bind(c, name='nekoamd_chkp_write')
       import c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: path, fields
       integer(c_int) :: tstep, nfields
       real(c_double) :: time, dt
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_chkp_read(path, tstep, time, dt, nfields, fields, n) &
          bind(c, name='nekoamd_chkp_read')
       import c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: path, fields
       integer(c_int) :: tstep, nfields
       real(c_double) :: time, dt
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_dissipation_objective(out, u, v, w, chi, bm, x, y, z, dx, &
          lx, nelv, k, obj_scale, if_lube) bind(c, name='nekoamd_dissipation_objective')
       import c_int, c_ptr, c_double
       real(c_double) :: out(3)
       type(c_ptr), value :: u, v, w, chi, bm, x, y, z, dx
       integer(c_int) :: lx, nelv, if_lube
       real(c_double) :: k, obj_scale
     end function

     integer(c_int) function nekoamd_dissipation_sensitivity(s, u, v, w, ua, va, wa, k, &
          obj_scale, if_lube, n) bind(c, name='nekoamd_dissipation_sensitivity')
       import c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: s, u, v, w, ua, va, wa
       real(c_double) :: k, obj_scale
       integer(c_int) :: if_lube
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_volume_constraint(out, rho, bm, volume_domain, v_max, &
          is_max, n) bind(c, name='nekoamd_volume_constraint')
       import c_int, c_ptr, c_double, c_int64_t
       real(c_double) :: out(2)
       type(c_ptr), value :: rho, bm
       real(c_double) :: volume_domain, v_max
       integer(c_int) :: is_max
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_volume_sensitivity(s, volume_domain, is_max, n) &
          bind(c, name='nekoamd_volume_sensitivity')
       import c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: s
       real(c_double) :: volume_domain
       integer(c_int) :: is_max
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_mma_create(mma, n, m, x, a0, a, c, d, xmin, xmax, &
          max_iter, epsimin, asyinit, asyincr, asydecr) bind(c, name='nekoamd_mma_create')
       import c_int, c_ptr, c_double
       type(c_ptr) :: mma
       integer(c_int) :: n, m, max_iter
       type(c_ptr), value :: x, xmin, xmax
       real(c_double) :: a0, a(*), c(*), d(*), epsimin, asyinit, asyincr, asydecr
     end function

     integer(c_int) function nekoamd_mma_free(mma) bind(c, name='nekoamd_mma_free')
       import c_int, c_ptr
       type(c_ptr), value :: mma
     end function

     integer(c_int) function nekoamd_mma_update(mma, iter, x, df0dx, fval, dfdx) &
          bind(c, name='nekoamd_mma_update')
       import c_int, c_ptr, c_double
       type(c_ptr), value :: mma, x, df0dx, dfdx
       integer(c_int) :: iter
       real(c_double) :: fval(*)
     end function

     integer(c_int) function nekoamd_mma_kkt(mma, x, df0dx, fval, dfdx) &
          bind(c, name='nekoamd_mma_kkt')
       import c_int, c_ptr, c_double
       type(c_ptr), value :: mma, x, df0dx, dfdx
       real(c_double) :: fval(*)
     end function

     integer(c_int) function nekoamd_mma_residuals(mma, res) bind(c, name='nekoamd_mma_residuals')
       import c_int, c_ptr, c_double
       type(c_ptr), value :: mma
       real(c_double) :: res(3)
     end function

     integer(c_int) function nekoamd_advection(fx, fy, fz, u, v, w, b, x, y, z, dx, lx, nelv) &
          bind(c, name='nekoamd_advection')
       import c_int, c_ptr
       type(c_ptr), value :: fx, fy, fz, u, v, w, b, x, y, z, dx
       integer(c_int) :: lx, nelv
     end function

     integer(c_int) function nekoamd_curl(cx, cy, cz, u, v, w, x, y, z, dx, lx, nelv) &
          bind(c, name='nekoamd_curl')
       import c_int, c_ptr
       type(c_ptr), value :: cx, cy, cz, u, v, w, x, y, z, dx
       integer(c_int) :: lx, nelv
     end function

     integer(c_int) function nekoamd_opgrad(ox, oy, oz, p, x, y, z, dx, wq, lx, nelv) &
          bind(c, name='nekoamd_opgrad')
       import c_int, c_ptr
       type(c_ptr), value :: ox, oy, oz, p, x, y, z, dx, wq
       integer(c_int) :: lx, nelv
     end function

     integer(c_int) function nekoamd_cdtp(out, tx, ty, tz, x, y, z, dx, wq, lx, nelv) &
          bind(c, name='nekoamd_cdtp')
       import c_int, c_ptr
       type(c_ptr), value :: out, tx, ty, tz, x, y, z, dx, wq
       integer(c_int) :: lx, nelv
     end function

     integer(c_int) function nekoamd_pnpn_prs_div(out, tx, ty, tz, binv, b, c, s, x, y, &
          z, dx, wgll, lx, nelv) bind(c, name='nekoamd_pnpn_prs_div')
       import c_int, c_ptr, c_double
       type(c_ptr), value :: out, tx, ty, tz, binv, b, c, x, y, z, dx, wgll
       real(c_double) :: s
       integer(c_int) :: lx, nelv
     end function

     integer(c_int) function nekoamd_surface_flux(out, u, v, w, x, y, z, dx, wq, bits, &
          lx, nelv) bind(c, name='nekoamd_surface_flux')
       import c_int, c_ptr
       type(c_ptr), value :: out, u, v, w, x, y, z, dx, wq, bits
       integer(c_int) :: lx, nelv
     end function

     integer(c_int) function nekoamd_box_face_bits(nx, ny, nz, faces, bits) &
          bind(c, name='nekoamd_box_face_bits')
       import c_int
       integer(c_int) :: nx, ny, nz, faces
       integer(c_int) :: bits(*)
     end function

     integer(c_int) function nekoamd_pnpn_extrapolate(ue, ve, we, u, v, w, ul, vl, wl, &
          ext, n) bind(c, name='nekoamd_pnpn_extrapolate')
       import c_int, c_ptr, c_int64_t, c_double
       type(c_ptr), value :: ue, ve, we, u, v, w, ul, vl, wl
       real(c_double) :: ext(2)
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_pnpn_makef(f_out, flag, ulag, f, u, b, coef, n) &
          bind(c, name='nekoamd_pnpn_makef')
       import c_int, c_ptr, c_int64_t, c_double
       type(c_ptr), value :: f_out, flag, ulag, f, u, b
       real(c_double) :: coef(4)
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_pnpn_extrapolate3(ue, ve, we, u, v, w, ul, vl, wl, &
          ul2, vl2, wl2, ext, n) bind(c, name='nekoamd_pnpn_extrapolate3')
       import c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: ue, ve, we, u, v, w, ul, vl, wl, ul2, vl2, wl2
       real(c_double) :: ext(3)
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_pnpn_makef3(f_out, flag, flag2, ulag, ulag2, f, u, b, &
          coef, n) bind(c, name='nekoamd_pnpn_makef3')
       import c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: f_out, flag, flag2, ulag, ulag2, f, u, b
       real(c_double) :: coef(6)
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_pnpn_makef_vector(f1o, f2o, f3o, fl1, fl2, fl3, &
          ul1, ul2, ul3, f1, f2, f3, u1, u2, u3, b, coef, n) &
          bind(c, name='nekoamd_pnpn_makef_vector')
       import c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: f1o, f2o, f3o, fl1, fl2, fl3, ul1, ul2, ul3, f1, f2, f3, &
            u1, u2, u3, b
       real(c_double) :: coef(4)
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_pnpn_prs_force(t, f, c, ue, chi, b, mu, rho, n) &
          bind(c, name='nekoamd_pnpn_prs_force')
       import c_int, c_ptr, c_int64_t, c_double
       type(c_ptr), value :: t, f, c, ue, chi, b
       real(c_double) :: mu, rho
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_pnpn_prs_force_vector(t1, t2, t3, f1, f2, f3, &
          c1, c2, c3, ue1, ue2, ue3, chi, b, mu, rho, n) &
          bind(c, name='nekoamd_pnpn_prs_force_vector')
       import c_int, c_ptr, c_int64_t, c_double
       type(c_ptr), value :: t1, t2, t3, f1, f2, f3, c1, c2, c3, ue1, ue2, ue3, chi, b
       real(c_double) :: mu, rho
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_brinkman_mass(fx, fy, fz, u, v, w, chi, b, n) &
          bind(c, name='nekoamd_brinkman_mass')
       import c_int, c_ptr, c_int64_t
       type(c_ptr), value :: fx, fy, fz, u, v, w, chi, b
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_sub2s(a, b, c, s, n) bind(c, name='nekoamd_sub2s')
       import c_int, c_ptr, c_int64_t, c_double
       type(c_ptr), value :: a, b, c
       real(c_double) :: s
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_pnpn_vel_res(r, f, g, n) bind(c, name='nekoamd_pnpn_vel_res')
       import c_int, c_ptr, c_int64_t
       type(c_ptr), value :: r, f, g
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_box_rng_uniform(out, nx, ny, nz, lx, seed) &
          bind(c, name='nekoamd_box_rng_uniform')
       import c_int, c_ptr, c_int64_t
       type(c_ptr), value :: out
       integer(c_int) :: nx, ny, nz, lx
       integer(c_int64_t) :: seed
     end function

     integer(c_int) function nekoamd_ramp_fwd(out, x, f_min, f_max, q, convex_up, n) &
          bind(c, name='nekoamd_ramp_fwd')
       import c_int, c_ptr, c_int64_t, c_double
       type(c_ptr), value :: out, x
       real(c_double) :: f_min, f_max, q
       integer(c_int) :: convex_up
       integer(c_int64_t) :: n
     end function

     integer(c_int) function nekoamd_ramp_bwd(df_din, df_dout, x, f_min, f_max, q, &
          convex_up, n) bind(c, name='nekoamd_ramp_bwd')
       import c_int, c_ptr, c_int64_t, c_double
       type(c_ptr), value :: df_din, df_dout, x
       real(c_double) :: f_min, f_max, q
       integer(c_int) :: convex_up
       integer(c_int64_t) :: n
     end function
  end interface

contains

  !> Abort with the library message when a C-ABI call failed.
  subroutine nk_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(len=512) :: msg
    if (rc .ne. 0) then
       call c_string(nekoamd_error_string(), msg)
       write(*, '(A,A,A,I0,A,A)') 'nekoamd error in ', where, ' (status ', rc, '): ', &
            trim(msg)
       error stop 'nekoamd: device call failed'
    end if
  end subroutine nk_check

  subroutine c_string(p, s)
    type(c_ptr), intent(in) :: p
    character(len=*), intent(out) :: s
    character(kind=c_char), pointer :: cs(:)
    integer :: i
    s = ''
    if (.not. c_associated(p)) return
    call c_f_pointer(p, cs, [len(s)])
    do i = 1, len(s)
       if (cs(i) .eq. c_null_char) exit
       s(i:i) = cs(i)
    end do
  end subroutine c_string

end module nekoamd_capi
