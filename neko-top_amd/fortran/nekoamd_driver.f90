!> bind(c) entry points of the Fortran host (declared in
!! include/nekoamd_fortran.h).  Each builds the synthetic box case through
!! the Neko-style types (space_t, box_mesh_t, gs_t, coef_t, bc_list_t) and
!! drives the plugin surface (ax_helm_factory, precon_factory,
!! krylov_solver_factory, set_pc, solve) exactly as the reference's solvers
!! call it (sources/adjoint/adjoint_scheme.f90:1046-1092,
!! sources/adjoint/adjoint_pnpn.f90:729-736).
module nekoamd_driver
  use, intrinsic :: iso_c_binding
  use nekoamd_capi
  use nekoamd_types
  use nekoamd_solvers
  implicit none
  private

  logical, save :: device_ready = .false.

contains

  subroutine ensure_device()
    if (.not. device_ready) then
       call nk_device_init(0)
       device_ready = .true.
    end if
  end subroutine ensure_device

  !> Jacobi-PCG solve on an nx*ny*nz box of order lx-1.
  !! kind 0: pressure Poisson (h1 = 1, Dirichlet at the outflow x = L);
  !! kind 1: velocity Helmholtz (h1 = 1/Re, h2 = 1.5/dt; Dirichlet at the
  !! inflow and the walls), with Re and dt given.
  !! f: right-hand side (gs-assembled and masked inside), x: solution;
  !! mon(1) = iterations, res(1:2) = start / final residual.
  function nekoamd_f_pcg_solve(nx, ny, nz, lx, jitter, seed, kind, re, dt, f, x, &
       abstol, max_iter, iters, res) result(rc) bind(c, name='nekoamd_f_pcg_solve')
    integer(c_int) :: nx, ny, nz, lx, kind, max_iter
    real(c_double) :: jitter, re, dt, abstol
    integer(c_int64_t) :: seed
    real(c_double), target :: f(*), x(*)
    integer(c_int) :: iters
    real(c_double) :: res(2)
    integer(c_int) :: rc
    type(space_t), target :: Xh
    type(box_mesh_t) :: msh
    type(gs_t), target :: gs_h
    type(coef_t), target :: coef
    type(bc_list_t) :: blst
    type(field_t) :: fd, xd
    class(ax_t), allocatable :: Ax
    class(pc_t), allocatable, target :: pc
    class(ksp_t), allocatable :: ksp
    type(ksp_monitor_t) :: mon
    integer(c_int64_t), allocatable :: gid(:)
    integer(c_int64_t) :: n

    call ensure_device()
    call Xh%init(lx)
    call msh%init(nx, ny, nz, jitter, seed)
    n = int(msh%nelv, c_int64_t) * lx**3
    allocate(gid(n))
    call nk_check(nekoamd_box_gids(msh%nx, msh%ny, msh%nz, Xh%lx, gid), 'box_gids')
    call gs_h%init(gid, n)
    call coef%init(Xh, msh, gs_h)
    call blst%init()
    if (kind .eq. 0) then
       call blst%add_faces(msh, lx, int(NK_FACE_XMAX))
       call coef%h1%fill(1.0_rp)
       coef%ifh2 = .false.
    else
       call blst%add_faces(msh, lx, int(ior(NK_FACE_XMIN, ior(NK_FACE_YMIN, &
            ior(NK_FACE_YMAX, ior(NK_FACE_ZMIN, NK_FACE_ZMAX))))))
       call coef%h1%fill(1.0_rp / re)
       call coef%h2%fill(1.5_rp / dt)
       coef%ifh2 = .true.
    end if
    call fd%init(n, 'rhs')
    call fd%upload(f)
    call gs_h%op(fd, GS_OP_ADD)
    call blst%apply_scalar(fd%x_d, n)
    call xd%init(n, 'x')

    call ax_helm_factory(Ax, full_formulation=.false.)
    call precon_factory(pc, 'jacobi')
    select type (p => pc)
    type is (jacobi_hip_t)
       call p%init(coef, n, gs_h)
    end select
    call krylov_solver_factory(ksp, n, 'cg', int(max_iter), abstol, coef)
    call ksp%set_pc(pc)
    call pc%update()
    mon = ksp%solve(Ax, xd, fd, n, coef, blst, gs_h)
    call nk_download(c_loc(x), xd%x_d, 8_c_int64_t * n)
    iters = mon%iter
    res(1) = mon%res_start
    res(2) = mon%res_final

    select type (k => ksp)
    type is (cg_hip_t)
       call k%free()
    end select
    select type (p => pc)
    type is (jacobi_hip_t)
       call p%free()
    end select
    call fd%free()
    call xd%free()
    call blst%free()
    call coef%free()
    call gs_h%free()
    call Xh%free()
    rc = 0
  end function nekoamd_f_pcg_solve

end module nekoamd_driver
