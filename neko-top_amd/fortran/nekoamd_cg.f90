!> The plugin surface the reference's solves call, backed by the HIP kernels:
!!
!!   ax_t        compute(w, u, coef, msh, Xh)        ax_helm_factory
!!                 (sources/adjoint/adjoint_pnpn.f90:196-217,
!!                  sources/mapping_functions/PDE_filter_mapping.f90:177)
!!   pc_t        update(), solve(z, r, n)            precon_factory
!!                 (sources/adjoint/adjoint_scheme.f90:1058-1092)
!!   ksp_t       solve(Ax, x, f, n, coef, blst, gs_h) -> ksp_monitor_t,
!!               solve_coupled(...) -> ksp_monitor_t(3),
!!               set_pc(pc)                          krylov_solver_factory
!!                 (adjoint_scheme.f90:1046-1055, :1090;
!!                  adjoint_pnpn.f90:734-736, :784-786)
!!
!! Fields are device buffers (field_t%x_d).  The CG for an ax_helm operator
!! with a Jacobi or identity preconditioner runs the fused device-resident
!! iteration of libnekoamd (scalars never leave HBM; the host polls the stop
!! flag every `poll` iterations); any other operator or preconditioner runs
!! the textbook loop below with host-side inner products.
!!
!! Solver types: 'cg' (FP64, the reference's) and 'cg_mixed' (FP32
!! iterations with FP64 reliable updates, stopping on the FP64 residual).
module nekoamd_solvers
  use, intrinsic :: iso_c_binding
  use nekoamd_capi
  use nekoamd_types
  implicit none
  private

  public :: ax_t, ax_helm_hip_t, ax_helm_factory, pc_t, jacobi_hip_t, ident_t, &
       hsmg_hip_t, precon_factory, ksp_t, cg_hip_t, gmres_hip_t, fcg_hip_t, ksp_monitor_t, &
       krylov_solver_factory, &
       projection_t, learned_t, LEARNED_OFF, LEARNED_HARVEST, LEARNED_APPLY, &
       LEARNED_APPLY_HARVEST

  type :: ksp_monitor_t
     integer :: iter = 0
     real(rp) :: res_start = 0.0_rp
     real(rp) :: res_final = 0.0_rp
  end type ksp_monitor_t

  ! ------------------------------------------------------------------ ax_t --
  type, abstract :: ax_t
   contains
     procedure(ax_compute_i), pass(this), deferred :: compute
     procedure, pass(this) :: compute_vector => ax_compute_vector
  end type ax_t

  !> Neko's ax_t%compute(w, u, coef, msh, Xh): the operator reads the
  !! element count from msh and the GLL derivative matrix from Xh.
  abstract interface
     subroutine ax_compute_i(this, w, u, coef, msh, Xh)
       import ax_t, field_t, coef_t, box_mesh_t, space_t
       class(ax_t), intent(in) :: this
       type(field_t), intent(inout) :: w
       type(field_t), intent(in) :: u
       type(coef_t), intent(in) :: coef
       type(box_mesh_t), intent(in) :: msh
       type(space_t), intent(in) :: Xh
     end subroutine ax_compute_i
  end interface

  !> Matrix-free Helmholtz operator w = A u with coef%h1, coef%h2, coef%ifh2.
  type, extends(ax_t) :: ax_helm_hip_t
   contains
     procedure, pass(this) :: compute => ax_helm_hip_compute
     procedure, pass(this) :: compute_vector => ax_helm_hip_compute_vector
  end type ax_helm_hip_t

  ! ------------------------------------------------------------------ pc_t --
  type, abstract :: pc_t
   contains
     procedure(pc_solve_i), pass(this), deferred :: solve
     procedure(pc_update_i), pass(this), deferred :: update
  end type pc_t

  abstract interface
     subroutine pc_solve_i(this, z, r, n)
       import pc_t, field_t, c_int64_t
       class(pc_t), intent(inout) :: this
       type(field_t), intent(inout) :: z
       type(field_t), intent(in) :: r
       integer(c_int64_t), intent(in) :: n
     end subroutine pc_solve_i
     subroutine pc_update_i(this)
       import pc_t
       class(pc_t), intent(inout) :: this
     end subroutine pc_update_i
  end interface

  !> Jacobi: d = 1 / gs(diag(A_L)) without the G_ab cross terms.
  type, extends(pc_t) :: jacobi_hip_t
     type(field_t) :: d
     type(coef_t), pointer :: coef => null()
     type(gs_t), pointer :: gs_h => null()
   contains
     procedure, pass(this) :: init => jacobi_init
     procedure, pass(this) :: solve => jacobi_solve
     procedure, pass(this) :: update => jacobi_update
     procedure, pass(this) :: free => jacobi_free
  end type jacobi_hip_t

  !> Identity preconditioner (the filter's 'ident').
  type, extends(pc_t) :: ident_t
     type(field_t) :: ones
   contains
     procedure, pass(this) :: solve => ident_solve
     procedure, pass(this) :: update => ident_update
  end type ident_t

  !> Hybrid Schwarz multigrid (Neko pc_hsmg, precon_factory 'hsmg', the
  !! pressure preconditioner of rugby_ball.case:46-52): levels lx, 4, 2;
  !! overlapping Schwarz with fast-diagonalisation local solves, coarse PCG.
  type, extends(pc_t) :: hsmg_hip_t
     type(c_ptr) :: h = c_null_ptr
   contains
     procedure, pass(this) :: init => hsmg_init
     procedure, pass(this) :: solve => hsmg_solve
     procedure, pass(this) :: update => hsmg_update
     procedure, pass(this) :: free => hsmg_free
  end type hsmg_hip_t

  ! ----------------------------------------------------------------- ksp_t --
  type, abstract :: ksp_t
     class(pc_t), pointer :: M => null()
     real(rp) :: abs_tol = 1e-8_rp
     integer :: max_iter = 800
   contains
     procedure, pass(this) :: set_pc => ksp_set_pc
     procedure(ksp_solve_i), pass(this), deferred :: solve
     procedure, pass(this) :: solve_coupled => ksp_solve_coupled
  end type ksp_t

  abstract interface
     function ksp_solve_i(this, Ax, x, f, n, coef, blst, gs_h) result(res)
       import ksp_t, ax_t, field_t, coef_t, bc_list_t, gs_t, ksp_monitor_t, c_int64_t
       class(ksp_t), intent(inout) :: this
       class(ax_t), intent(in) :: Ax
       type(field_t), intent(inout) :: x
       type(field_t), intent(in) :: f
       integer(c_int64_t), intent(in) :: n
       type(coef_t), intent(inout) :: coef
       type(bc_list_t), intent(inout) :: blst
       type(gs_t), intent(inout) :: gs_h
       type(ksp_monitor_t) :: res
     end function ksp_solve_i
  end interface

  !> Jacobi/identity preconditioned conjugate gradients.
  type, extends(ksp_t) :: cg_hip_t
     type(c_ptr) :: state = c_null_ptr
     integer(c_int) :: precision = NK_CG_FP64
     !> work vectors of the generic loop (allocated on first use)
     type(field_t) :: r, p, w, z
     integer :: poll = 8
     !> reliable updates of the last mixed-precision solve
     integer :: nupdate = 0
     !> the second and third component's states of a coupled solve
     type(c_ptr) :: state3(2) = c_null_ptr
     integer(c_int64_t) :: n = 0
     integer(c_int) :: nelv = 0
   contains
     procedure, pass(this) :: init => cg_init
     procedure, pass(this) :: solve => cg_solve
     procedure, pass(this) :: solve_coupled => cg_solve_coupled
     procedure, pass(this) :: free => cg_free
  end type cg_hip_t

  !> Restarted flexible GMRES (Neko gmres_t, krylov_solver_factory
  !! 'gmres'), restart length lgmres (Neko's default 30); jacobi, ident or
  !! hsmg preconditioner.
  type, extends(ksp_t) :: gmres_hip_t
     type(c_ptr) :: state = c_null_ptr
     integer(c_int) :: lgmres = 30
   contains
     procedure, pass(this) :: init => gmres_init
     procedure, pass(this) :: solve => gmres_solve
     procedure, pass(this) :: free => gmres_free
  end type gmres_hip_t

  !> Flexible PCG (Polak-Ribiere beta) with the Jacobi preconditioner and,
  !! when mlp is set, the element network's Galerkin-weighted correction
  !! every `every`-th iteration (libnekoamd's nekoamd_fcg_solve): the host of
  !! the learned coarse correction.  With mlp null the iterates are PCG's.
  !! No reference counterpart ('fcg' is this build's solver type).
  type, extends(ksp_t) :: fcg_hip_t
     type(c_ptr) :: state = c_null_ptr
     type(c_ptr) :: mlp = c_null_ptr
     integer(c_int) :: every = 1
     integer(c_int) :: poll = 8
   contains
     procedure, pass(this) :: init => fcg_init
     procedure, pass(this) :: solve => fcg_solve
     procedure, pass(this) :: free => fcg_free
  end type fcg_hip_t

  ! ------------------------------------------------------------ projection --
  !> Neko's projection_t as the reference drives it (proj_prs, proj_u/v/w:
  !! sources/adjoint/adjoint_pnpn.f90:102-105, init :365-373, pre_solving
  !! :729-730, :777-779, post_solving :739-740, :788-793): up to L previous
  !! solutions, A-orthogonal, give the initial guess of the next solve.
  !! Active from time step activ_step + 1 on; L = 0 switches it off.  keep:
  !! solutions whose span survives the restart of a full basis (0, Neko's
  !! restart, keeps none).
  type :: projection_t
     type(c_ptr) :: h = c_null_ptr
     integer :: L = 0
     integer :: keep = 0
     integer :: activ_step = 5
   contains
     procedure, pass(this) :: init => proj_init
     procedure, pass(this) :: free => proj_free
     procedure, pass(this) :: clear => proj_clear
     procedure, pass(this) :: active => proj_active
     procedure, pass(this) :: pre_solving => proj_pre_solving
     procedure, pass(this) :: post_solving => proj_post_solving
     procedure, pass(this) :: set_fp32 => proj_set_fp32
  end type projection_t

  ! --------------------------------------------------------------- learned --
  integer, parameter :: LEARNED_OFF = 0, LEARNED_HARVEST = 1, LEARNED_APPLY = 2, &
       LEARNED_APPLY_HARVEST = 3

  !> The learned initial-guess direction of a solve (libnekoamd's element
  !! network): pre() before ksp%solve (after the projection's pre_solving),
  !! post() after it (before the projection's post_solving).  Harvest mode
  !! stores the (residual, correction) pair of every solve for training.
  !! No reference counterpart.
  type :: learned_t
     type(c_ptr) :: h = c_null_ptr
     integer :: mode = LEARNED_OFF
     !> > 0: the network's correction inside a flexible-CG pressure solve
     !! every correct_every-th iteration (needs solver 'fcg')
     integer :: correct_every = 0
     real(rp) :: gamma = 0.0_rp
   contains
     procedure, pass(this) :: init => learned_init
     procedure, pass(this) :: free => learned_free
     procedure, pass(this) :: pre => learned_pre
     procedure, pass(this) :: post => learned_post
     procedure, pass(this) :: train => learned_train
     procedure, pass(this) :: pairs => learned_pairs
  end type learned_t

contains

  ! ------------------------------------------------------------ factories --
  subroutine ax_helm_factory(Ax, full_formulation)
    class(ax_t), allocatable, intent(inout) :: Ax
    logical, intent(in) :: full_formulation
    if (allocated(Ax)) deallocate(Ax)
    if (full_formulation) error stop 'ax_helm_factory: full-stress form is not provided'
    allocate(ax_helm_hip_t :: Ax)
  end subroutine ax_helm_factory

  subroutine precon_factory(pc, pctype)
    class(pc_t), allocatable, intent(inout) :: pc
    character(len=*), intent(in) :: pctype
    if (allocated(pc)) deallocate(pc)
    select case (trim(pctype))
    case ('jacobi')
       allocate(jacobi_hip_t :: pc)
    case ('ident')
       allocate(ident_t :: pc)
    case ('hsmg')
       allocate(hsmg_hip_t :: pc)
    case default
       write(*, '(A,A)') 'precon_factory: unknown preconditioner ', trim(pctype)
       error stop 'precon_factory: unknown preconditioner'
    end select
  end subroutine precon_factory

  subroutine krylov_solver_factory(ksp, n, solver, max_iter, abstol, coef)
    class(ksp_t), allocatable, intent(inout) :: ksp
    integer(c_int64_t), intent(in) :: n
    character(len=*), intent(in) :: solver
    integer, intent(in) :: max_iter
    real(rp), intent(in) :: abstol
    type(coef_t), intent(in) :: coef
    if (allocated(ksp)) deallocate(ksp)
    select case (trim(solver))
    case ('cg', 'cg_mixed')
       allocate(cg_hip_t :: ksp)
    case ('gmres')
       allocate(gmres_hip_t :: ksp)
    case ('fcg')
       allocate(fcg_hip_t :: ksp)
    case default
       write(*, '(A,A)') 'krylov_solver_factory: unknown solver ', trim(solver)
       error stop 'krylov_solver_factory: unknown solver'
    end select
    ksp%max_iter = max_iter
    ksp%abs_tol = abstol
    select type (k => ksp)
    type is (cg_hip_t)
       if (trim(solver) .eq. 'cg_mixed') then
          call k%init(n, coef%nelv, NK_CG_MIXED)
       else
          call k%init(n, coef%nelv, NK_CG_FP64)
       end if
    type is (gmres_hip_t)
       call k%init(n, 30)
    type is (fcg_hip_t)
       call k%init(n, coef%nelv)
    end select
  end subroutine krylov_solver_factory

  ! ------------------------------------------------------------------ hsmg --
  !> faces: the Dirichlet box faces of the preconditioned operator;
  !! precision: NK_CG_FP64, or NK_CG_MIXED for an FP32 data path
  subroutine hsmg_init(this, msh, lx, faces, precision)
    class(hsmg_hip_t), intent(inout) :: this
    type(box_mesh_t), intent(in) :: msh
    integer, intent(in) :: lx, faces, precision
    integer(c_int) :: nx, ny, nz, l, f, crs, prec
    call this%free()
    nx = msh%nx
    ny = msh%ny
    nz = msh%nz
    l = lx
    f = faces
    crs = 10
    prec = precision
    call nk_check(nekoamd_hsmg_create(this%h, nx, ny, nz, l, msh%corners, f, crs, prec), &
         'hsmg_create')
  end subroutine hsmg_init

  subroutine hsmg_solve(this, z, r, n)
    class(hsmg_hip_t), intent(inout) :: this
    type(field_t), intent(inout) :: z
    type(field_t), intent(in) :: r
    integer(c_int64_t), intent(in) :: n
    call nk_check(nekoamd_hsmg_apply(this%h, z%x_d, r%x_d), 'hsmg_apply')
  end subroutine hsmg_solve

  !> the levels depend on the mesh only: nothing to update
  subroutine hsmg_update(this)
    class(hsmg_hip_t), intent(inout) :: this
  end subroutine hsmg_update

  subroutine hsmg_free(this)
    class(hsmg_hip_t), intent(inout) :: this
    if (c_associated(this%h)) call nk_check(nekoamd_hsmg_free(this%h), 'hsmg_free')
    this%h = c_null_ptr
  end subroutine hsmg_free

  ! ----------------------------------------------------------------- gmres --
  subroutine gmres_init(this, n, m)
    class(gmres_hip_t), intent(inout) :: this
    integer(c_int64_t), intent(in) :: n
    integer, intent(in) :: m
    integer(c_int64_t) :: nn
    call this%free()
    nn = n
    this%lgmres = int(m, c_int)
    call nk_check(nekoamd_gmres_create(this%state, nn, this%lgmres), 'gmres_create')
  end subroutine gmres_init

  subroutine gmres_free(this)
    class(gmres_hip_t), intent(inout) :: this
    if (c_associated(this%state)) call nk_check(nekoamd_gmres_free(this%state), &
         'gmres_free')
    this%state = c_null_ptr
  end subroutine gmres_free

  function gmres_solve(this, Ax, x, f, n, coef, blst, gs_h) result(res)
    class(gmres_hip_t), intent(inout) :: this
    class(ax_t), intent(in) :: Ax
    type(field_t), intent(inout) :: x
    type(field_t), intent(in) :: f
    integer(c_int64_t), intent(in) :: n
    type(coef_t), intent(inout) :: coef
    type(bc_list_t), intent(inout) :: blst
    type(gs_t), intent(inout) :: gs_h
    type(ksp_monitor_t) :: res
    type(nk_cg_monitor_c) :: mon
    type(c_ptr) :: pc
    integer(c_int) :: kind, mi, lx, ifh2
    integer(c_int64_t) :: nmask
    real(c_double) :: vol, tol
    if (.not. associated(this%M)) error stop 'gmres_solve: no preconditioner set (set_pc)'
    if (n .ne. coef%n) error stop 'gmres_solve: n does not match coef'
    select type (A => Ax)
    type is (ax_helm_hip_t)
    class default
       error stop 'gmres_solve: the device GMRES applies ax_helm'
    end select
    pc = c_null_ptr
    select type (P => this%M)
    type is (jacobi_hip_t)
       kind = 1
       pc = P%d%x_d
    type is (ident_t)
       kind = 0
    type is (hsmg_hip_t)
       kind = 2
       pc = P%h
    class default
       error stop 'gmres_solve: unsupported preconditioner'
    end select
    vol = coef%volume
    tol = this%abs_tol
    mi = this%max_iter
    lx = coef%lx
    ifh2 = merge(1, 0, coef%ifh2)
    nmask = blst%nall
    call nk_check(nekoamd_gmres_solve(this%state, x%x_d, f%x_d, kind, pc, coef%mult%x_d, &
         vol, tol, mi, gs_h%h, blst%all_d, nmask, coef%Xh%dx_d, coef%h1%x_d, coef%h2%x_d, &
         coef%G11%x_d, coef%G12%x_d, coef%G13%x_d, coef%G22%x_d, coef%G23%x_d, &
         coef%G33%x_d, coef%B%x_d, lx, ifh2, mon), 'gmres_solve')
    res%iter = mon%iter
    res%res_start = mon%res_start
    res%res_final = mon%res_final
  end function gmres_solve

  ! ------------------------------------------------------------------- fcg --
  subroutine fcg_init(this, n, nelv)
    class(fcg_hip_t), intent(inout) :: this
    integer(c_int64_t), intent(in) :: n
    integer, intent(in) :: nelv
    integer(c_int64_t) :: nn
    integer(c_int) :: ne
    call this%free()
    nn = n
    ne = nelv
    call nk_check(nekoamd_fcg_create(this%state, nn, ne), 'fcg_create')
  end subroutine fcg_init

  subroutine fcg_free(this)
    class(fcg_hip_t), intent(inout) :: this
    if (c_associated(this%state)) call nk_check(nekoamd_fcg_free(this%state), 'fcg_free')
    this%state = c_null_ptr
    this%mlp = c_null_ptr
  end subroutine fcg_free

  !> x = A^{-1} f, x0 = 0, for ax_helm with the Jacobi preconditioner.
  function fcg_solve(this, Ax, x, f, n, coef, blst, gs_h) result(res)
    class(fcg_hip_t), intent(inout) :: this
    class(ax_t), intent(in) :: Ax
    type(field_t), intent(inout) :: x
    type(field_t), intent(in) :: f
    integer(c_int64_t), intent(in) :: n
    type(coef_t), intent(inout) :: coef
    type(bc_list_t), intent(inout) :: blst
    type(gs_t), intent(inout) :: gs_h
    type(ksp_monitor_t) :: res
    type(nk_cg_monitor_c) :: mon
    type(c_ptr) :: d_d
    integer(c_int) :: mi, lx, ifh2
    integer(c_int64_t) :: nmask
    real(c_double) :: vol, tol
    if (.not. associated(this%M)) error stop 'fcg_solve: no preconditioner set (set_pc)'
    if (n .ne. coef%n) error stop 'fcg_solve: n does not match coef'
    select type (A => Ax)
    type is (ax_helm_hip_t)
    class default
       error stop 'fcg_solve: the device FCG applies ax_helm'
    end select
    select type (P => this%M)
    type is (jacobi_hip_t)
       d_d = P%d%x_d
    class default
       error stop 'fcg_solve: the device FCG takes the jacobi preconditioner'
    end select
    vol = coef%volume
    tol = this%abs_tol
    mi = this%max_iter
    lx = coef%lx
    ifh2 = merge(1, 0, coef%ifh2)
    nmask = blst%nall
    call nk_check(nekoamd_fcg_solve(this%state, x%x_d, f%x_d, d_d, coef%mult%x_d, vol, tol, &
         mi, this%poll, gs_h%h, blst%all_d, nmask, coef%Xh%dx_d, coef%h1%x_d, coef%h2%x_d, &
         coef%G11%x_d, coef%G12%x_d, coef%G13%x_d, coef%G22%x_d, coef%G23%x_d, &
         coef%G33%x_d, coef%B%x_d, lx, ifh2, this%mlp, this%every, mon), 'fcg_solve')
    res%iter = mon%iter
    res%res_start = mon%res_start
    res%res_final = mon%res_final
  end function fcg_solve

  ! ------------------------------------------------------------ projection --
  subroutine proj_init(this, n, L, activ_step, keep)
    class(projection_t), intent(inout) :: this
    integer(c_int64_t), intent(in) :: n
    integer, intent(in) :: L
    integer, intent(in), optional :: activ_step, keep
    integer(c_int64_t) :: nn
    integer(c_int) :: ll, kk
    call this%free()
    this%L = L
    this%keep = 0
    if (present(activ_step)) this%activ_step = activ_step
    if (L .le. 0) return
    nn = n
    ll = L
    call nk_check(nekoamd_proj_create(this%h, nn, ll), 'proj_create')
    if (present(keep)) then
       if (keep .gt. 0) then
          kk = keep
          call nk_check(nekoamd_proj_set_keep(this%h, kk), 'proj_set_keep')
          this%keep = keep
       end if
    end if
  end subroutine proj_init

  !> The basis in FP32 (nekoamd_proj_set_fp32): half the bytes per solve;
  !! pre applies the operator of the last post_solving.  Call while the basis
  !! is empty (after init).
  subroutine proj_set_fp32(this, on)
    class(projection_t), intent(inout) :: this
    logical, intent(in) :: on
    integer(c_int) :: f
    if (.not. c_associated(this%h)) return
    f = merge(1, 0, on)
    call nk_check(nekoamd_proj_set_fp32(this%h, f), 'proj_set_fp32')
  end subroutine proj_set_fp32

  subroutine proj_free(this)
    class(projection_t), intent(inout) :: this
    if (c_associated(this%h)) call nk_check(nekoamd_proj_free(this%h), 'proj_free')
    this%h = c_null_ptr
    this%L = 0
  end subroutine proj_free

  subroutine proj_clear(this)
    class(projection_t), intent(inout) :: this
    if (c_associated(this%h)) call nk_check(nekoamd_proj_clear(this%h), 'proj_clear')
  end subroutine proj_clear

  logical function proj_active(this, tstep)
    class(projection_t), intent(in) :: this
    integer, intent(in) :: tstep
    proj_active = this%L .gt. 0 .and. tstep .gt. this%activ_step .and. &
         c_associated(this%h)
  end function proj_active

  !> b <- b - A xbar with xbar the projection of the solution on the basis.
  subroutine proj_pre_solving(this, b, tstep, coef, n)
    class(projection_t), intent(inout) :: this
    type(field_t), intent(inout) :: b
    integer, intent(in) :: tstep
    type(coef_t), intent(in) :: coef
    integer(c_int64_t), intent(in) :: n
    if (.not. this%active(tstep)) return
    if (n .ne. coef%n) error stop 'projection: n does not match coef'
    call nk_check(nekoamd_proj_pre(this%h, b%x_d, coef%mult%x_d), 'proj_pre')
  end subroutine proj_pre_solving

  !> x <- x + xbar; x joins the basis with its image under the solve's
  !! operator (coef%h1, coef%h2, coef%ifh2 as the solve used them).
  subroutine proj_post_solving(this, x, Ax, coef, blst, gs_h, n, tstep)
    class(projection_t), intent(inout) :: this
    type(field_t), intent(inout) :: x
    class(ax_t), intent(in) :: Ax
    type(coef_t), intent(in) :: coef
    type(bc_list_t), intent(in) :: blst
    type(gs_t), intent(in) :: gs_h
    integer(c_int64_t), intent(in) :: n
    integer, intent(in) :: tstep
    integer(c_int) :: lx, nelv, ifh2
    integer(c_int64_t) :: nmask
    if (.not. this%active(tstep)) return
    select type (A => Ax)
    type is (ax_helm_hip_t)
    class default
       error stop 'projection: post_solving needs the ax_helm operator'
    end select
    lx = coef%lx
    nelv = coef%nelv
    ifh2 = merge(1, 0, coef%ifh2)
    nmask = blst%nall
    call nk_check(nekoamd_proj_post(this%h, x%x_d, coef%mult%x_d, gs_h%h, blst%all_d, &
         nmask, coef%Xh%dx_d, coef%h1%x_d, coef%h2%x_d, coef%G11%x_d, coef%G12%x_d, &
         coef%G13%x_d, coef%G22%x_d, coef%G23%x_d, coef%G33%x_d, coef%B%x_d, lx, nelv, &
         ifh2), 'proj_post')
  end subroutine proj_post_solving

  ! --------------------------------------------------------------- learned --
  subroutine learned_init(this, nx, ny, nz, lx, capacity, batch, seed)
    class(learned_t), intent(inout) :: this
    integer, intent(in) :: nx, ny, nz, lx, capacity, batch
    integer(c_int64_t), intent(in) :: seed
    integer(c_int) :: a, b, c, l, cap, bt
    call this%free()
    a = nx
    b = ny
    c = nz
    l = lx
    cap = capacity
    bt = batch
    call nk_check(nekoamd_mlp_create(this%h, a, b, c, l, cap, bt, seed), 'mlp_create')
  end subroutine learned_init

  subroutine learned_free(this)
    class(learned_t), intent(inout) :: this
    if (c_associated(this%h)) call nk_check(nekoamd_mlp_free(this%h), 'mlp_free')
    this%h = c_null_ptr
    this%mode = LEARNED_OFF
  end subroutine learned_free

  !> b <- b - gamma A q with the network's direction q (apply modes); keeps
  !! b for the harvest (harvest modes).  The operator is the solve's (coef).
  subroutine learned_pre(this, b, coef, blst, gs_h)
    class(learned_t), intent(inout) :: this
    type(field_t), intent(inout) :: b
    type(coef_t), intent(in) :: coef
    type(bc_list_t), intent(in) :: blst
    type(gs_t), intent(in) :: gs_h
    integer(c_int) :: lx, ifh2, apply, save
    integer(c_int64_t) :: nmask
    real(c_double) :: g
    this%gamma = 0.0_rp
    if (this%mode .eq. LEARNED_OFF .or. .not. c_associated(this%h)) return
    lx = coef%lx
    ifh2 = merge(1, 0, coef%ifh2)
    apply = merge(1, 0, this%mode .ge. LEARNED_APPLY)
    save = merge(1, 0, this%mode .eq. LEARNED_HARVEST .or. &
         this%mode .eq. LEARNED_APPLY_HARVEST)
    nmask = blst%nall
    call nk_check(nekoamd_mlp_pre(this%h, b%x_d, coef%mult%x_d, gs_h%h, blst%all_d, nmask, &
         coef%Xh%dx_d, coef%h1%x_d, coef%h2%x_d, coef%G11%x_d, coef%G12%x_d, coef%G13%x_d, &
         coef%G22%x_d, coef%G23%x_d, coef%G33%x_d, coef%B%x_d, lx, ifh2, apply, save, g), &
         'mlp_pre')
    this%gamma = g
  end subroutine learned_pre

  !> x <- x + gamma q; harvest the pair (kept b, x) in harvest modes.
  subroutine learned_post(this, x)
    class(learned_t), intent(inout) :: this
    type(field_t), intent(inout) :: x
    integer(c_int) :: hv
    if (this%mode .eq. LEARNED_OFF .or. .not. c_associated(this%h)) return
    hv = merge(1, 0, this%mode .eq. LEARNED_HARVEST .or. &
         this%mode .eq. LEARNED_APPLY_HARVEST)
    call nk_check(nekoamd_mlp_post(this%h, x%x_d, hv), 'mlp_post')
  end subroutine learned_post

  subroutine learned_train(this, epochs, batch, lr, seed, loss0, loss1)
    class(learned_t), intent(inout) :: this
    integer, intent(in) :: epochs, batch
    real(c_float), intent(in) :: lr
    integer(c_int64_t), intent(in) :: seed
    real(c_double), intent(out) :: loss0, loss1
    integer(c_int) :: e, b
    real(c_float) :: l
    integer(c_int64_t) :: s
    e = epochs
    b = batch
    l = lr
    s = seed
    if (.not. c_associated(this%h)) error stop 'learned: train before init'
    call nk_check(nekoamd_mlp_train(this%h, e, b, l, s, loss0, loss1), 'mlp_train')
  end subroutine learned_train

  integer function learned_pairs(this)
    class(learned_t), intent(in) :: this
    integer(c_int) :: n
    n = 0
    if (c_associated(this%h)) call nk_check(nekoamd_mlp_pairs(this%h, n), 'mlp_pairs')
    learned_pairs = n
  end function learned_pairs

  ! -------------------------------------------------------------------- ax --
  subroutine ax_helm_hip_compute(this, w, u, coef, msh, Xh)
    class(ax_helm_hip_t), intent(in) :: this
    type(field_t), intent(inout) :: w
    type(field_t), intent(in) :: u
    type(coef_t), intent(in) :: coef
    type(box_mesh_t), intent(in) :: msh
    type(space_t), intent(in) :: Xh
    integer(c_int) :: lx, nelv, ifh2
    lx = Xh%lx
    nelv = msh%nelv
    if (lx .ne. coef%lx .or. nelv .ne. coef%nelv) &
         error stop 'ax_t%compute: msh / Xh do not match coef'
    ifh2 = merge(1, 0, coef%ifh2)
    call nk_check(nekoamd_ax_helm_f64(w%x_d, u%x_d, Xh%dx_d, coef%h1%x_d, &
         coef%h2%x_d, coef%G11%x_d, coef%G12%x_d, coef%G13%x_d, coef%G22%x_d, &
         coef%G23%x_d, coef%G33%x_d, coef%B%x_d, lx, nelv, ifh2), 'ax_helm')
  end subroutine ax_helm_hip_compute

  !> Neko's ax_t%compute_vector(au, av, aw, u, v, w, coef, msh, Xh): the
  !! operator on the three velocity components (the velocity residual).
  !! Default: three compute calls.
  subroutine ax_compute_vector(this, au, av, aw, u, v, w, coef, msh, Xh)
    class(ax_t), intent(in) :: this
    type(field_t), intent(inout) :: au, av, aw
    type(field_t), intent(in) :: u, v, w
    type(coef_t), intent(in) :: coef
    type(box_mesh_t), intent(in) :: msh
    type(space_t), intent(in) :: Xh
    call this%compute(au, u, coef, msh, Xh)
    call this%compute(av, v, coef, msh, Xh)
    call this%compute(aw, w, coef, msh, Xh)
  end subroutine ax_compute_vector

  !> The three components in one launch: the operator data are read once,
  !! each result equals compute on that component.
  subroutine ax_helm_hip_compute_vector(this, au, av, aw, u, v, w, coef, msh, Xh)
    class(ax_helm_hip_t), intent(in) :: this
    type(field_t), intent(inout) :: au, av, aw
    type(field_t), intent(in) :: u, v, w
    type(coef_t), intent(in) :: coef
    type(box_mesh_t), intent(in) :: msh
    type(space_t), intent(in) :: Xh
    integer(c_int) :: lx, nelv, ifh2
    lx = Xh%lx
    nelv = msh%nelv
    if (lx .ne. coef%lx .or. nelv .ne. coef%nelv) &
         error stop 'ax_t%compute_vector: msh / Xh do not match coef'
    ifh2 = merge(1, 0, coef%ifh2)
    call nk_check(nekoamd_ax_helm_vector_f64(au%x_d, av%x_d, aw%x_d, u%x_d, v%x_d, &
         w%x_d, Xh%dx_d, coef%h1%x_d, coef%h2%x_d, coef%G11%x_d, coef%G12%x_d, &
         coef%G13%x_d, coef%G22%x_d, coef%G23%x_d, coef%G33%x_d, coef%B%x_d, lx, &
         nelv, ifh2), 'ax_helm_vector')
  end subroutine ax_helm_hip_compute_vector

  ! ---------------------------------------------------------------- jacobi --
  !> jacobi_t%init(coef, dof, gs) (sources/adjoint/adjoint_scheme.f90:1071-1076)
  subroutine jacobi_init(this, coef, n, gs_h)
    class(jacobi_hip_t), intent(inout) :: this
    type(coef_t), intent(in), target :: coef
    integer(c_int64_t), intent(in) :: n
    type(gs_t), intent(in), target :: gs_h
    this%coef => coef
    this%gs_h => gs_h
    call this%d%init(n, 'jacobi_d')
    call this%update()
  end subroutine jacobi_init

  subroutine jacobi_update(this)
    class(jacobi_hip_t), intent(inout) :: this
    integer(c_int) :: lx, nelv, ifh2
    associate(c => this%coef)
      lx = c%lx
      nelv = c%nelv
      ifh2 = merge(1, 0, c%ifh2)
      call nk_check(nekoamd_jacobi_update(this%d%x_d, this%gs_h%h, c%Xh%dx_d, &
           c%G11%x_d, c%G22%x_d, c%G33%x_d, c%h1%x_d, c%h2%x_d, c%B%x_d, lx, nelv, &
           ifh2), 'jacobi_update')
    end associate
  end subroutine jacobi_update

  subroutine jacobi_solve(this, z, r, n)
    class(jacobi_hip_t), intent(inout) :: this
    type(field_t), intent(inout) :: z
    type(field_t), intent(in) :: r
    integer(c_int64_t), intent(in) :: n
    integer(c_int64_t) :: nn
    nn = n
    call nk_check(nekoamd_jacobi_solve(z%x_d, r%x_d, this%d%x_d, nn), 'jacobi_solve')
  end subroutine jacobi_solve

  subroutine jacobi_free(this)
    class(jacobi_hip_t), intent(inout) :: this
    call this%d%free()
    nullify(this%coef, this%gs_h)
  end subroutine jacobi_free

  subroutine ident_solve(this, z, r, n)
    class(ident_t), intent(inout) :: this
    type(field_t), intent(inout) :: z
    type(field_t), intent(in) :: r
    integer(c_int64_t), intent(in) :: n
    integer(c_int64_t) :: nn
    nn = n
    call nk_check(nekoamd_copy(z%x_d, r%x_d, nn), 'copy')
  end subroutine ident_solve

  subroutine ident_update(this)
    class(ident_t), intent(inout) :: this
  end subroutine ident_update

  ! ------------------------------------------------------------------- ksp --
  subroutine ksp_set_pc(this, pc)
    class(ksp_t), intent(inout) :: this
    class(pc_t), intent(in), target :: pc
    this%M => pc
  end subroutine ksp_set_pc

  !> Three independent solves with one operator, as the reference's
  !! ksp_vel%solve_coupled (sources/adjoint/adjoint_pnpn.f90:784-786).
  function ksp_solve_coupled(this, Ax, x, y, z, fx, fy, fz, n, coef, blx, bly, blz, &
       gs_h) result(res)
    class(ksp_t), intent(inout) :: this
    class(ax_t), intent(in) :: Ax
    type(field_t), intent(inout) :: x, y, z
    type(field_t), intent(in) :: fx, fy, fz
    integer(c_int64_t), intent(in) :: n
    type(coef_t), intent(inout) :: coef
    type(bc_list_t), intent(inout) :: blx, bly, blz
    type(gs_t), intent(inout) :: gs_h
    type(ksp_monitor_t) :: res(3)
    res(1) = this%solve(Ax, x, fx, n, coef, blx, gs_h)
    res(2) = this%solve(Ax, y, fy, n, coef, bly, gs_h)
    res(3) = this%solve(Ax, z, fz, n, coef, blz, gs_h)
  end function ksp_solve_coupled

  subroutine cg_init(this, n, nelv, precision)
    class(cg_hip_t), intent(inout) :: this
    integer(c_int64_t), intent(in) :: n
    integer(c_int), intent(in) :: nelv, precision
    integer(c_int64_t) :: nn
    integer(c_int) :: ne, pr
    call this%free()
    nn = n
    ne = nelv
    pr = precision
    this%precision = precision
    this%n = n
    this%nelv = nelv
    call nk_check(nekoamd_cg_create(this%state, nn, ne, pr), 'cg_create')
  end subroutine cg_init

  subroutine cg_free(this)
    class(cg_hip_t), intent(inout) :: this
    if (c_associated(this%state)) call nk_check(nekoamd_cg_free(this%state), 'cg_free')
    this%state = c_null_ptr
    if (c_associated(this%state3(1))) call nk_check(nekoamd_cg_free(this%state3(1)), 'cg_free')
    if (c_associated(this%state3(2))) call nk_check(nekoamd_cg_free(this%state3(2)), 'cg_free')
    this%state3 = c_null_ptr
    call this%r%free()
    call this%p%free()
    call this%w%free()
    call this%z%free()
  end subroutine cg_free

  !> solve_coupled (adjoint_pnpn.f90:784-786).  Mixed precision with
  !! ax_helm, a Jacobi preconditioner and one mask for the three components:
  !! the three solves run together (nekoamd_cg_solve3: one set-up, one
  !! driver); otherwise three solves one after the other.
  function cg_solve_coupled(this, Ax, x, y, z, fx, fy, fz, n, coef, blx, bly, blz, &
       gs_h) result(res)
    class(cg_hip_t), intent(inout) :: this
    class(ax_t), intent(in) :: Ax
    type(field_t), intent(inout) :: x, y, z
    type(field_t), intent(in) :: fx, fy, fz
    integer(c_int64_t), intent(in) :: n
    type(coef_t), intent(inout) :: coef
    type(bc_list_t), intent(inout) :: blx, bly, blz
    type(gs_t), intent(inout) :: gs_h
    type(ksp_monitor_t) :: res(3)
    type(nk_cg_monitor_c) :: mon(3)
    type(c_ptr) :: st(3), xs(3), fs(3), d_d
    logical :: fused, same_mask
    real(c_double) :: vol, tol
    integer(c_int) :: mi, poll, lx, ifh2, ne, pr
    integer(c_int64_t) :: nmask, nn
    integer :: c

    fused = .false.
    d_d = c_null_ptr
    same_mask = blx%nall .eq. bly%nall .and. blx%nall .eq. blz%nall
    if (same_mask .and. blx%nall .gt. 0) same_mask = &
         c_associated(blx%all_d, bly%all_d) .and. c_associated(blx%all_d, blz%all_d)
    if (this%precision .eq. NK_CG_MIXED .and. associated(this%M) .and. same_mask &
         .and. n .eq. coef%n .and. n .eq. this%n) then
       select type (A => Ax)
       type is (ax_helm_hip_t)
          select type (P => this%M)
          type is (jacobi_hip_t)
             fused = .true.
             d_d = P%d%x_d
          end select
       end select
    end if
    if (.not. fused) then
       res = ksp_solve_coupled(this, Ax, x, y, z, fx, fy, fz, n, coef, blx, bly, blz, gs_h)
       return
    end if
    nn = this%n
    ne = this%nelv
    pr = this%precision
    do c = 1, 2
       if (.not. c_associated(this%state3(c))) &
            call nk_check(nekoamd_cg_create(this%state3(c), nn, ne, pr), 'cg_create')
    end do
    st = [this%state, this%state3(1), this%state3(2)]
    xs = [x%x_d, y%x_d, z%x_d]
    fs = [fx%x_d, fy%x_d, fz%x_d]
    vol = coef%volume
    tol = this%abs_tol
    mi = this%max_iter
    poll = this%poll
    lx = coef%lx
    ifh2 = merge(1, 0, coef%ifh2)
    nmask = blx%nall
    call nk_check(nekoamd_cg_solve3(st, xs, fs, d_d, coef%mult%x_d, vol, tol, mi, poll, &
         gs_h%h, blx%all_d, nmask, coef%Xh%dx_d, coef%h1%x_d, coef%h2%x_d, coef%G11%x_d, &
         coef%G12%x_d, coef%G13%x_d, coef%G22%x_d, coef%G23%x_d, coef%G33%x_d, &
         coef%B%x_d, lx, ifh2, mon), 'cg_solve3')
    this%nupdate = 0
    do c = 1, 3
       res(c)%iter = mon(c)%iter
       res(c)%res_start = mon(c)%res_start
       res(c)%res_final = mon(c)%res_final
       this%nupdate = this%nupdate + mon(c)%nupdate
    end do
  end function cg_solve_coupled

  !> x = A^{-1} f, x0 = 0; f gs-assembled and zero on blst.
  function cg_solve(this, Ax, x, f, n, coef, blst, gs_h) result(res)
    class(cg_hip_t), intent(inout) :: this
    class(ax_t), intent(in) :: Ax
    type(field_t), intent(inout) :: x
    type(field_t), intent(in) :: f
    integer(c_int64_t), intent(in) :: n
    type(coef_t), intent(inout) :: coef
    type(bc_list_t), intent(inout) :: blst
    type(gs_t), intent(inout) :: gs_h
    type(ksp_monitor_t) :: res
    type(nk_cg_monitor_c) :: mon
    type(c_ptr) :: d_d
    logical :: fused
    real(c_double) :: vol, tol
    integer(c_int) :: mi, poll, lx, ifh2
    integer(c_int64_t) :: nmask

    if (.not. associated(this%M)) error stop 'cg_solve: no preconditioner set (set_pc)'
    if (n .ne. coef%n) error stop 'cg_solve: n does not match coef'
    fused = .false.
    d_d = c_null_ptr
    select type (A => Ax)
    type is (ax_helm_hip_t)
       select type (P => this%M)
       type is (jacobi_hip_t)
          fused = .true.
          d_d = P%d%x_d
       type is (ident_t)
          if (P%ones%n .ne. n) then
             call P%ones%init(n)
             call P%ones%fill(1.0_rp)
          end if
          fused = .true.
          d_d = P%ones%x_d
       end select
    end select

    if (fused) then
       vol = coef%volume
       tol = this%abs_tol
       mi = this%max_iter
       poll = this%poll
       lx = coef%lx
       ifh2 = merge(1, 0, coef%ifh2)
       nmask = blst%nall
       call nk_check(nekoamd_cg_solve(this%state, x%x_d, f%x_d, d_d, coef%mult%x_d, &
            vol, tol, mi, poll, gs_h%h, &
            blst%all_d, nmask, coef%Xh%dx_d, coef%h1%x_d, coef%h2%x_d, coef%G11%x_d, &
            coef%G12%x_d, coef%G13%x_d, coef%G22%x_d, coef%G23%x_d, coef%G33%x_d, &
            coef%B%x_d, lx, ifh2, mon), 'cg_solve')
       res%iter = mon%iter
       res%res_start = mon%res_start
       res%res_final = mon%res_final
       this%nupdate = mon%nupdate
    else
       if (this%precision .ne. NK_CG_FP64) error stop &
            'cg_solve: cg_mixed needs ax_helm with a jacobi or ident preconditioner'
       res = cg_solve_generic(this, Ax, x, f, n, coef, blst, gs_h)
    end if
  end function cg_solve

  !> Textbook PCG for any ax_t / pc_t (host-side inner products).
  function cg_solve_generic(this, Ax, x, f, n, coef, blst, gs_h) result(res)
    class(cg_hip_t), intent(inout) :: this
    class(ax_t), intent(in) :: Ax
    type(field_t), intent(inout) :: x
    type(field_t), intent(in) :: f
    integer(c_int64_t), intent(in) :: n
    type(coef_t), intent(inout) :: coef
    type(bc_list_t), intent(inout) :: blst
    type(gs_t), intent(inout) :: gs_h
    type(ksp_monitor_t) :: res
    real(c_double) :: rtr, rtz1, rtz2, pap, alpha, beta, rnorm, norm_fac, a
    integer :: iter
    if (this%r%n .ne. n) call this%r%init(n, 'cg_r')
    if (this%p%n .ne. n) call this%p%init(n, 'cg_p')
    if (this%w%n .ne. n) call this%w%init(n, 'cg_w')
    if (this%z%n .ne. n) call this%z%init(n, 'cg_z')
    norm_fac = 1.0_rp / sqrt(coef%volume)
    call nk_check(nekoamd_rzero(x%x_d, n), 'rzero')
    call nk_check(nekoamd_rzero(this%p%x_d, n), 'rzero')
    call nk_check(nekoamd_copy(this%r%x_d, f%x_d, n), 'copy')
    call nk_check(nekoamd_glsc3(rtr, this%r%x_d, coef%mult%x_d, this%r%x_d, n), 'glsc3')
    rnorm = sqrt(rtr) * norm_fac
    res%res_start = rnorm
    res%res_final = rnorm
    res%iter = 0
    if (rnorm .eq. 0.0_rp) return
    rtz1 = 1.0_rp
    do iter = 1, this%max_iter
       call this%M%solve(this%z, this%r, n)
       rtz2 = rtz1
       call nk_check(nekoamd_glsc3(rtz1, this%r%x_d, coef%mult%x_d, this%z%x_d, n), &
            'glsc3')
       beta = rtz1 / rtz2
       if (iter .eq. 1) beta = 0.0_rp
       ! p = z + beta p
       call nk_check(nekoamd_cmult(this%p%x_d, beta, n), 'cmult')
       call nk_check(nekoamd_add2(this%p%x_d, this%z%x_d, n), 'add2')
       call Ax%compute(this%w, this%p, coef, coef%msh, coef%Xh)
       call gs_h%op(this%w, GS_OP_ADD)
       call blst%apply_scalar(this%w%x_d, n)
       call nk_check(nekoamd_glsc3(pap, this%w%x_d, coef%mult%x_d, this%p%x_d, n), &
            'glsc3')
       alpha = rtz1 / pap
       a = -alpha
       call nk_check(nekoamd_add2s2(x%x_d, this%p%x_d, alpha, n), 'add2s2')
       call nk_check(nekoamd_add2s2(this%r%x_d, this%w%x_d, a, n), 'add2s2')
       call nk_check(nekoamd_glsc3(rtr, this%r%x_d, coef%mult%x_d, this%r%x_d, n), &
            'glsc3')
       rnorm = sqrt(rtr) * norm_fac
       res%iter = iter
       if (rnorm .lt. this%abs_tol) exit
    end do
    res%res_final = rnorm
  end function cg_solve_generic

end module nekoamd_solvers
