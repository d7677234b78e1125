!> One Pn/Pn time step of the Brinkman-penalised channel on the device, in
!! the order of sources/adjoint/adjoint_pnpn.f90:619-827:
!!   sumab (u_e)                                           :665-666
!!   source terms / advection into the explicit forcing    :669-682
!!   makeabf, makebdf, lag update                          :688-700
!!   pressure residual, gs, bc                             :716-726
!!   proj_prs%pre_solving                                  :729-730
!!   pc%update, ksp_prs%solve, p += dp                     :732-743
!!   proj_prs%post_solving                                 :739-740
!!   velocity residual, gs, bc                             :747-773
!!   proj_u/v/w%pre_solving                                :777-779
!!   pc%update, ksp_vel%solve_coupled, u += du             :781-801
!!   proj_u/v/w%post_solving                               :788-793
!! The residuals restate Neko's pnpn_prs_res / pnpn_vel_res (not vendored):
!!   A_p dp = cdtp(Binv gs(F - mu B curl curl u_e - chi B u_e)) - A_p p
!!            - (rho bd0/dt) int_Gamma_D q u.n
!!   H du   = F - opgrad(p) - H u,   H = mu A + (rho bd0/dt + chi) B
!! with F = EXT(B f) + (B rho/dt)(bd1 u + bd2 u_lag), f = -(u.grad)u plus
!! the explicit source terms of the handler (source_term%compute).  The
!! Brinkman term is a source term of the handler; by default it is
!! semi-implicit (chi, its implicit sink, goes into the Helmholtz h2 and the
!! pressure forcing: the reference's explicit treatment is unstable at
!! chi dt = 10); set_brinkman(implicit=.false.) gives the reference's
!! explicit treatment (f -= chi u, then EXT), stable for chi dt < 4/3.
!!
!! Adjoint mode (make_adjoint): the state is the adjoint velocity and
!! pressure about a frozen base flow (ub, vb, wb), the step being
!! adjoint_pnpn_step itself: the forcing f = obj_scale curl(curl ub) +
!! K chi ub (adjoint_pnpn.f90:669 with the adjoint minimum-dissipation and
!! lube source terms) times B, then the adjoint advection about ub
!! (:676-678); everything after is the same step.
module nekoamd_fluid
  use, intrinsic :: iso_c_binding
  use nekoamd_capi
  use nekoamd_types
  use nekoamd_solvers
  use nekoamd_advection
  use nekoamd_source_term
  implicit none
  private

  public :: fluid_pnpn_t

  integer, parameter, public :: NK_VEL_FACES = &
       ior(NK_FACE_XMIN, ior(NK_FACE_YMIN, ior(NK_FACE_YMAX, ior(NK_FACE_ZMIN, NK_FACE_ZMAX))))
  integer, parameter, public :: NK_WALL_FACES = &
       ior(NK_FACE_YMIN, ior(NK_FACE_YMAX, ior(NK_FACE_ZMIN, NK_FACE_ZMAX)))

  type :: fluid_pnpn_t
     type(space_t) :: Xh
     type(box_mesh_t) :: msh
     type(gs_t) :: gs_h
     type(coef_t) :: c_Xh
     type(field_t) :: u, v, w, p, ulag, vlag, wlag, fxlag, fylag, fzlag
     !> the second lags of time_order 3 (allocated only then)
     type(field_t) :: ulag2, vlag2, wlag2, fxlag2, fylag2, fzlag2
     type(field_t) :: ue, ve, we, fx, fy, fz, Fbx, Fby, Fbz
     type(field_t) :: t1, t2, t3, c1, c2, c3, p_res, dp, du, dv, dw
     type(field_t) :: chi, Binv, h1v, h2v
     type(c_ptr) :: face_bits_d = c_null_ptr
     type(bc_list_t) :: bclst_vel, bclst_prs, bclst_walls
     class(ax_t), allocatable :: Ax
     class(ksp_t), allocatable :: ksp_prs, ksp_vel
     class(pc_t), allocatable :: pc_prs, pc_vel
     type(projection_t) :: proj_prs, proj_u, proj_v, proj_w
     type(learned_t) :: learned_prs
     logical :: adjoint = .false.
     type(field_t) :: ub, vb, wb
     type(advection_adjoint_t) :: adv
     !> the momentum source terms (the Brinkman term by default)
     type(source_term_handler_t) :: source_term
     type(simple_brinkman_source_term_t) :: brinkman
     real(rp) :: K = 1.0_rp, obj_scale = 1.0_rp
     real(rp) :: gamma_prs = 0.0_rp
     real(rp) :: Re = 100.0_rp, dt = 1e-3_rp, rho = 1.0_rp, mu = 0.01_rp
     real(rp) :: bd_vel = -1.0_rp       ! bd0 the velocity Jacobi was built with
     logical :: pc_prs_ready = .false.
     integer :: tstep = 0
     !> case.numerics.time_order: 2 (BDF2/EXT2) or 3 (BDF3/EXT3)
     integer :: time_order = 2
     integer :: iters(4) = 0
     real(rp) :: res(4) = 0.0_rp
   contains
     procedure, pass(this) :: init => fluid_init
     procedure, pass(this) :: step => fluid_step
     procedure, pass(this) :: free => fluid_free
     procedure, pass(this) :: projection => fluid_projection
     procedure, pass(this) :: make_adjoint => fluid_make_adjoint
     procedure, pass(this) :: reset => fluid_reset
     procedure, pass(this) :: set_dealias => fluid_set_dealias
     procedure, pass(this) :: set_pressure_solver => fluid_set_pressure_solver
     procedure, pass(this) :: set_brinkman => fluid_set_brinkman
     procedure, pass(this) :: set_time_order => fluid_set_time_order
     procedure, pass(this) :: checkpoint => fluid_checkpoint
     procedure, pass(this) :: restart => fluid_restart
  end type fluid_pnpn_t

contains

  !> BDF/EXT coefficients of step tstep for a constant step, ramped as
  !! Neko's ext_bdf does: order min(tstep, time_order).  bd(1) u^{n+1} =
  !! sum_j bd(j+1) u^{n+1-j} + dt (...), u^{n+1} ~ sum_j ext(j) u^{n+1-j}:
  !!   order 1: bd = 1, 1;  ext = 1
  !!   order 2: bd = 3/2, 2, -1/2;  ext = 2, -1
  !!   order 3: bd = 11/6, 3, -3/2, 1/3;  ext = 3, -3, 1
  subroutine coeffs(tstep, order, bd, ext)
    integer, intent(in) :: tstep, order
    real(rp), intent(out) :: bd(4), ext(3)
    bd = 0.0_rp
    ext = 0.0_rp
    select case (min(tstep, order))
    case (:1)
       bd(1:2) = [1.0_rp, 1.0_rp]
       ext(1) = 1.0_rp
    case (2)
       bd(1:3) = [1.5_rp, 2.0_rp, -0.5_rp]
       ext(1:2) = [2.0_rp, -1.0_rp]
    case default
       bd = [11.0_rp / 6.0_rp, 3.0_rp, -1.5_rp, 1.0_rp / 3.0_rp]
       ext = [3.0_rp, -3.0_rp, 1.0_rp]
    end select
  end subroutine coeffs

  !> solver: krylov_solver_factory type of both solves ('cg' or 'cg_mixed');
  !! time_order: case.numerics.time_order -- 2 (BDF2/EXT2 after a BDF1/EXT1
  !! first step, the default) or 3 (then BDF3/EXT3 from the third step, as
  !! the reference's cases run, rugby_ball.case:12)
  subroutine fluid_init(this, nx, ny, nz, lx, jitter, seed, Re, dt, abstol_p, &
       abstol_v, max_iter, solver, time_order)
    class(fluid_pnpn_t), intent(inout), target :: this
    integer, intent(in) :: nx, ny, nz, lx, max_iter
    real(rp), intent(in) :: jitter, Re, dt, abstol_p, abstol_v
    character(len=*), intent(in) :: solver
    integer(c_int64_t), intent(in) :: seed
    integer, intent(in), optional :: time_order
    integer(c_int64_t), allocatable :: gid(:)
    integer(c_int), allocatable, target :: bits(:)
    integer(c_int64_t) :: n
    integer(c_int) :: fv
    this%time_order = 2
    if (present(time_order)) then
       if (time_order .ne. 2 .and. time_order .ne. 3) error stop &
            'fluid_pnpn_t: time_order must be 2 or 3'
       this%time_order = time_order
    end if
    this%Re = Re
    this%dt = dt
    this%mu = 1.0_rp / Re
    call this%Xh%init(lx)
    call this%msh%init(nx, ny, nz, jitter, seed)
    n = int(this%msh%nelv, c_int64_t) * lx**3
    allocate(gid(n))
    call nk_check(nekoamd_box_gids(this%msh%nx, this%msh%ny, this%msh%nz, this%Xh%lx, &
         gid), 'box_gids')
    call this%gs_h%init(gid, n)
    deallocate(gid)
    call this%c_Xh%init(this%Xh, this%msh, this%gs_h)
    call this%bclst_vel%init()
    call this%bclst_vel%add_faces(this%msh, lx, NK_VEL_FACES)
    call this%bclst_prs%init()
    call this%bclst_prs%add_faces(this%msh, lx, int(NK_FACE_XMAX))
    call this%bclst_walls%init()
    call this%bclst_walls%add_faces(this%msh, lx, NK_WALL_FACES)
    allocate(bits(this%msh%nelv))
    fv = NK_VEL_FACES
    call nk_check(nekoamd_box_face_bits(this%msh%nx, this%msh%ny, this%msh%nz, fv, bits), &
         'face_bits')
    call nk_alloc(this%face_bits_d, int(4, c_int64_t) * this%msh%nelv)
    call nk_upload(this%face_bits_d, c_loc(bits), int(4, c_int64_t) * this%msh%nelv)
    call this%u%init(n, 'u')
    call this%v%init(n, 'v')
    call this%w%init(n, 'w')
    call this%p%init(n, 'p')
    call this%ulag%init(n)
    call this%vlag%init(n)
    call this%wlag%init(n)
    call this%fxlag%init(n)
    call this%fylag%init(n)
    call this%fzlag%init(n)
    if (this%time_order .eq. 3) call init_lags2(this)
    call this%ue%init(n)
    call this%ve%init(n)
    call this%we%init(n)
    call this%fx%init(n)
    call this%fy%init(n)
    call this%fz%init(n)
    call this%Fbx%init(n)
    call this%Fby%init(n)
    call this%Fbz%init(n)
    call this%t1%init(n)
    call this%t2%init(n)
    call this%t3%init(n)
    call this%c1%init(n)
    call this%c2%init(n)
    call this%c3%init(n)
    call this%p_res%init(n)
    call this%dp%init(n)
    call this%du%init(n)
    call this%dv%init(n)
    call this%dw%init(n)
    call this%chi%init(n, 'chi')
    call this%h1v%init(n)
    call this%h2v%init(n)
    ! Binv = 1 / gs(B)
    call this%Binv%init(n, 'Binv')
    call nk_check(nekoamd_copy(this%Binv%x_d, this%c_Xh%B%x_d, n), 'copy')
    call this%gs_h%op(this%Binv, GS_OP_ADD)
    call nk_check(nekoamd_invcol1(this%Binv%x_d, n), 'invcol1')
    ! initial state: uniform inflow u = (1,0,0), no slip on the walls
    call this%u%fill(1.0_rp)
    call this%bclst_walls%apply_scalar(this%u%x_d, n)
    call nk_check(nekoamd_copy(this%ulag%x_d, this%u%x_d, n), 'copy')
    call reset_lags2(this)
    ! operators and solvers (adjoint_scheme.f90:693-727, :820-855,
    ! adjoint_pnpn.f90:196-217)
    call ax_helm_factory(this%Ax, full_formulation=.false.)
    call krylov_solver_factory(this%ksp_prs, n, solver, max_iter, abstol_p, this%c_Xh)
    call krylov_solver_factory(this%ksp_vel, n, solver, max_iter, abstol_v, this%c_Xh)
    call precon_factory(this%pc_prs, 'jacobi')
    call precon_factory(this%pc_vel, 'jacobi')
    select type (pp => this%pc_prs)
    type is (jacobi_hip_t)
       call pp%init(this%c_Xh, n, this%gs_h)
    end select
    select type (pv => this%pc_vel)
    type is (jacobi_hip_t)
       call pv%init(this%c_Xh, n, this%gs_h)
    end select
    call this%ksp_prs%set_pc(this%pc_prs)
    call this%ksp_vel%set_pc(this%pc_vel)
    ! collocated advection until set_dealias (advection_adjoint_fctry.f90:57-99)
    call advection_adjoint_factory(this%adv, this%c_Xh, .false., 0)
    ! the Brinkman term on the state velocity (steady_state_problem.f90:147-163)
    call this%source_term%init(n)
    call this%brinkman%init(this%fx, this%fy, this%fz, this%c_Xh, this%chi, this%u, &
         this%v, this%w, implicit=.true.)
    call this%source_term%add(this%brinkman)
    this%tstep = 0
    this%pc_prs_ready = .false.
    this%bd_vel = -1.0_rp
  end subroutine fluid_init

  !> Projection spaces of the pressure and velocity solves (the reference's
  !! pr_projection_dim / vel_projection_dim, activ_step; 0 = off).  keep_prs /
  !! keep_vel: solutions kept when a full basis restarts (default 0, Neko's
  !! restart).
  subroutine fluid_projection(this, L_prs, L_vel, activ_step, keep_prs, keep_vel)
    class(fluid_pnpn_t), intent(inout) :: this
    integer, intent(in) :: L_prs, L_vel, activ_step
    integer, intent(in), optional :: keep_prs, keep_vel
    integer(c_int64_t) :: n
    integer :: kp, kv
    n = this%c_Xh%n
    kp = 0
    kv = 0
    if (present(keep_prs)) kp = keep_prs
    if (present(keep_vel)) kv = keep_vel
    call this%proj_prs%init(n, L_prs, activ_step, kp)
    call this%proj_u%init(n, L_vel, activ_step, kv)
    call this%proj_v%init(n, L_vel, activ_step, kv)
    call this%proj_w%init(n, L_vel, activ_step, kv)
  end subroutine fluid_projection

  !> Turn the case into the adjoint solve: adjoint state zero, base flow
  !! fields ub, vb, wb (zero until set), forcing weights K and obj_scale.
  subroutine fluid_make_adjoint(this, kcoef, obj_scale)
    class(fluid_pnpn_t), intent(inout) :: this
    real(rp), intent(in) :: kcoef, obj_scale
    integer(c_int64_t) :: n
    n = this%c_Xh%n
    this%adjoint = .true.
    this%K = kcoef
    this%obj_scale = obj_scale
    if (this%ub%n .ne. n) then
       call this%ub%init(n, 'ub')
       call this%vb%init(n, 'vb')
       call this%wb%init(n, 'wb')
    end if
    call this%u%fill(0.0_rp)
    call this%v%fill(0.0_rp)
    call this%w%fill(0.0_rp)
    call this%p%fill(0.0_rp)
    call this%ulag%fill(0.0_rp)
    call this%vlag%fill(0.0_rp)
    call this%wlag%fill(0.0_rp)
    call this%fxlag%fill(0.0_rp)
    call this%fylag%fill(0.0_rp)
    call this%fzlag%fill(0.0_rp)
    call reset_lags2(this)
    this%tstep = 0
  end subroutine fluid_make_adjoint

  subroutine init_lags2(this)
    type(fluid_pnpn_t), intent(inout) :: this
    integer(c_int64_t) :: n
    n = this%c_Xh%n
    if (this%ulag2%n .eq. n) return
    call this%ulag2%init(n)
    call this%vlag2%init(n)
    call this%wlag2%init(n)
    call this%fxlag2%init(n)
    call this%fylag2%init(n)
    call this%fzlag2%init(n)
  end subroutine init_lags2

  !> second lags as the first ones start: ulag2 = u, the rest zero
  subroutine reset_lags2(this)
    type(fluid_pnpn_t), intent(inout) :: this
    integer(c_int64_t) :: n
    if (this%time_order .ne. 3) return
    n = this%c_Xh%n
    call nk_check(nekoamd_copy(this%ulag2%x_d, this%u%x_d, n), 'copy')
    call this%vlag2%fill(0.0_rp)
    call this%wlag2%fill(0.0_rp)
    call this%fxlag2%fill(0.0_rp)
    call this%fylag2%fill(0.0_rp)
    call this%fzlag2%fill(0.0_rp)
  end subroutine reset_lags2

  !> case.numerics.time_order of a fluid that has not stepped yet (2 or 3)
  subroutine fluid_set_time_order(this, order)
    class(fluid_pnpn_t), intent(inout) :: this
    integer, intent(in) :: order
    if (order .ne. 2 .and. order .ne. 3) error stop 'fluid_pnpn_t: time_order must be 2 or 3'
    if (this%tstep .ne. 0) error stop 'fluid_pnpn_t: set the time order before the first step'
    this%time_order = order
    if (order .eq. 3) then
       call init_lags2(this)
       call reset_lags2(this)
    end if
  end subroutine fluid_set_time_order

  subroutine free_prs_solver(this)
    type(fluid_pnpn_t), intent(inout) :: this
    if (allocated(this%ksp_prs)) then
       select type (k => this%ksp_prs)
       type is (cg_hip_t)
          call k%free()
       type is (gmres_hip_t)
          call k%free()
       end select
    end if
    if (allocated(this%pc_prs)) then
       select type (p => this%pc_prs)
       type is (jacobi_hip_t)
          call p%free()
       type is (hsmg_hip_t)
          call p%free()
       end select
    end if
  end subroutine free_prs_solver

  !> pressure_solver type and preconditioner (case.fluid.pressure_solver:
  !! 'cg' | 'cg_mixed' | 'gmres' with 'jacobi' | 'hsmg' | 'hsmg_mixed' (its
  !! data path in FP32), rugby_ball.case:46-52), keeping its tolerance and
  !! iteration cap
  subroutine fluid_set_pressure_solver(this, solver, pctype)
    class(fluid_pnpn_t), intent(inout), target :: this
    character(len=*), intent(in) :: solver, pctype
    real(rp) :: tol
    integer :: mi
    integer(c_int64_t) :: n
    n = this%c_Xh%n
    tol = this%ksp_prs%abs_tol
    mi = this%ksp_prs%max_iter
    call free_prs_solver(this)
    call krylov_solver_factory(this%ksp_prs, n, solver, mi, tol, this%c_Xh)
    if (trim(pctype) .eq. 'hsmg_mixed') then
       call precon_factory(this%pc_prs, 'hsmg')
    else
       call precon_factory(this%pc_prs, pctype)
    end if
    select type (pp => this%pc_prs)
    type is (jacobi_hip_t)
       call pp%init(this%c_Xh, n, this%gs_h)
    type is (hsmg_hip_t)
       call pp%init(this%msh, int(this%c_Xh%lx), int(NK_FACE_XMAX), &
            merge(int(NK_CG_MIXED), int(NK_CG_FP64), trim(pctype) .eq. 'hsmg_mixed'))
    end select
    call this%ksp_prs%set_pc(this%pc_prs)
    this%pc_prs_ready = .false.
  end subroutine fluid_set_pressure_solver

  !> The Brinkman term semi-implicit (default) or explicit, as the reference
  !! treats it (source_term%compute, opcolv, makeabf: adjoint_pnpn.f90:669-691)
  subroutine fluid_set_brinkman(this, implicit)
    class(fluid_pnpn_t), intent(inout) :: this
    logical, intent(in) :: implicit
    this%brinkman%implicit = implicit
    this%bd_vel = -1.0_rp
  end subroutine fluid_set_brinkman

  !> The fields a run restarts from: u, v, w, p, their velocity lags and the
  !! lagged explicit forcing (ten; with time_order 3 the six second lags
  !! too, sixteen).  nf: how many.
  subroutine state_fields(this, ptrs, nf)
    class(fluid_pnpn_t), intent(in) :: this
    type(c_ptr), intent(out) :: ptrs(16)
    integer(c_int), intent(out) :: nf
    ptrs = c_null_ptr
    ptrs(1:10) = [this%u%x_d, this%v%x_d, this%w%x_d, this%p%x_d, this%ulag%x_d, &
         this%vlag%x_d, this%wlag%x_d, this%fxlag%x_d, this%fylag%x_d, this%fzlag%x_d]
    nf = 10
    if (this%time_order .eq. 3) then
       ptrs(11:16) = [this%ulag2%x_d, this%vlag2%x_d, this%wlag2%x_d, this%fxlag2%x_d, &
            this%fylag2%x_d, this%fzlag2%x_d]
       nf = 16
    end if
  end subroutine state_fields

  !> Write the state and the step counter to a checkpoint file (chkp_t,
  !! sources/adjoint/adjoint_scheme.f90:159).
  subroutine fluid_checkpoint(this, path)
    class(fluid_pnpn_t), intent(inout) :: this
    character(len=*), intent(in) :: path
    character(len=:), allocatable, target :: cpath
    type(c_ptr), target :: ptrs(16)
    integer(c_int) :: ts, nf
    real(c_double) :: t, dt
    integer(c_int64_t) :: n
    cpath = trim(path) // c_null_char
    call state_fields(this, ptrs, nf)
    ts = int(this%tstep, c_int)
    t = this%tstep * this%dt
    dt = this%dt
    n = this%c_Xh%n
    call nk_check(nekoamd_chkp_write(c_loc(cpath), ts, t, dt, nf, c_loc(ptrs), n), 'chkp_write')
  end subroutine fluid_checkpoint

  !> Restart from a checkpoint as adjoint_pnpn_restart does
  !! (sources/adjoint/adjoint_pnpn.f90:418-484): read the state, then make
  !! u, v, w, p and the velocity lags continuous again by col2 with mult and
  !! gs ADD (the lagged forcing is left as read: its derivatives are not
  !! continuous across elements).  The next step is step tstep + 1.
  subroutine fluid_restart(this, path)
    class(fluid_pnpn_t), intent(inout) :: this
    character(len=*), intent(in) :: path
    character(len=:), allocatable, target :: cpath
    type(c_ptr), target :: ptrs(16)
    integer(c_int) :: ts, nf
    real(c_double) :: t, dt
    integer(c_int64_t) :: n
    integer :: i
    cpath = trim(path) // c_null_char
    call state_fields(this, ptrs, nf)
    n = this%c_Xh%n
    call nk_check(nekoamd_chkp_read(c_loc(cpath), ts, t, dt, nf, c_loc(ptrs), n), 'chkp_read')
    if (abs(dt - this%dt) > 1e-12_rp * abs(this%dt)) &
         error stop 'restart: the checkpoint was written with another time step'
    do i = 1, nf
       ! u, v, w, p and the velocity lags (not the lagged forcing)
       if (i .ge. 8 .and. i .le. 10) cycle
       if (i .ge. 14) cycle
       call nk_check(nekoamd_col2(ptrs(i), this%c_Xh%mult%x_d, n), 'col2')
       call this%gs_h%op_ptr(ptrs(i), n, GS_OP_ADD)
    end do
    this%tstep = int(ts)
  end subroutine fluid_restart

  !> case.numerics.dealias: the advection terms on lxd Gauss-Legendre
  !! points (lxd <= 0: 3 (p + 1) / 2) instead of the collocated forms
  subroutine fluid_set_dealias(this, dealias, lxd)
    class(fluid_pnpn_t), intent(inout), target :: this
    logical, intent(in) :: dealias
    integer, intent(in) :: lxd
    call advection_adjoint_factory(this%adv, this%c_Xh, dealias, lxd)
  end subroutine fluid_set_dealias

  !> Back to the initial state (the reference resets the case between
  !! design iterations, rugby_ball/driver.f90:133-137): forward cases
  !! restart from uniform inflow with no-slip walls, adjoint cases from rest;
  !! projection bases are cleared, the velocity Jacobi is rebuilt.
  subroutine fluid_reset(this)
    class(fluid_pnpn_t), intent(inout) :: this
    integer(c_int64_t) :: n
    n = this%c_Xh%n
    if (this%adjoint) then
       call this%u%fill(0.0_rp)
    else
       call this%u%fill(1.0_rp)
       call this%bclst_walls%apply_scalar(this%u%x_d, n)
    end if
    call nk_check(nekoamd_copy(this%ulag%x_d, this%u%x_d, n), 'copy')
    call this%v%fill(0.0_rp)
    call this%w%fill(0.0_rp)
    call this%p%fill(0.0_rp)
    call this%vlag%fill(0.0_rp)
    call this%wlag%fill(0.0_rp)
    call this%fxlag%fill(0.0_rp)
    call this%fylag%fill(0.0_rp)
    call this%fzlag%fill(0.0_rp)
    call reset_lags2(this)
    call this%proj_prs%clear()
    call this%proj_u%clear()
    call this%proj_v%clear()
    call this%proj_w%clear()
    this%tstep = 0
    this%bd_vel = -1.0_rp
  end subroutine fluid_reset

  !> gs-average a field with the mass: a = gs(B a) / gs(B) (col2, gs_op,
  !! col2 in one pass over the shared points)
  subroutine mass_average(this, a)
    type(fluid_pnpn_t), intent(inout) :: this
    type(field_t), intent(inout) :: a
    call nk_check(nekoamd_gs_mass_average(this%gs_h%h, a%x_d, this%c_Xh%B%x_d, &
         this%Binv%x_d), 'gs_mass_average')
  end subroutine mass_average

  subroutine fluid_step(this)
    class(fluid_pnpn_t), intent(inout), target :: this
    real(rp) :: bd(4), ext(3)
    real(c_double) :: mk(4), mk3(6), mu, rho, s, e2(2)
    real(rp) :: t
    integer(c_int) :: lx, nelv
    integer(c_int64_t) :: n
    type(ksp_monitor_t) :: mp, mv(3)
    type(c_ptr) :: sink

    this%tstep = this%tstep + 1
    t = this%tstep * this%dt
    call coeffs(this%tstep, this%time_order, bd, ext)
    associate(c => this%c_Xh, Xh => this%Xh)
      n = c%n
      lx = c%lx
      nelv = c%nelv
      mu = this%mu
      rho = this%rho
      ! (1) extrapolated velocity
      if (this%time_order .eq. 3) then
         call nk_check(nekoamd_pnpn_extrapolate3(this%ue%x_d, this%ve%x_d, this%we%x_d, &
              this%u%x_d, this%v%x_d, this%w%x_d, this%ulag%x_d, this%vlag%x_d, &
              this%wlag%x_d, this%ulag2%x_d, this%vlag2%x_d, this%wlag2%x_d, ext, n), &
              'extrapolate3')
      else
         e2 = ext(1:2)
         call nk_check(nekoamd_pnpn_extrapolate(this%ue%x_d, this%ve%x_d, this%we%x_d, &
              this%u%x_d, this%v%x_d, this%w%x_d, this%ulag%x_d, this%vlag%x_d, &
              this%wlag%x_d, e2, n), 'extrapolate')
      end if
      if (.not. this%adjoint) then
         ! (2) explicit forcing B f = -B (u.grad) u (adv%compute)
         call this%adv%compute(this%u, this%v, this%w, this%fx, this%fy, this%fz)
      else
         ! (2) adjoint source terms times B, then the adjoint advection
         call nk_check(nekoamd_curl(this%t1%x_d, this%t2%x_d, this%t3%x_d, this%ub%x_d, &
              this%vb%x_d, this%wb%x_d, c%x%x_d, c%y%x_d, c%z%x_d, Xh%dx_d, lx, nelv), &
              'curl')
         call mass_average(this, this%t1)
         call mass_average(this, this%t2)
         call mass_average(this, this%t3)
         call nk_check(nekoamd_curl(this%c1%x_d, this%c2%x_d, this%c3%x_d, this%t1%x_d, &
              this%t2%x_d, this%t3%x_d, c%x%x_d, c%y%x_d, c%z%x_d, Xh%dx_d, lx, nelv), &
              'curl')
         call mass_average(this, this%c1)
         call mass_average(this, this%c2)
         call mass_average(this, this%c3)
         call nk_check(nekoamd_adjoint_forcing(this%fx%x_d, this%fy%x_d, this%fz%x_d, &
              this%c1%x_d, this%c2%x_d, this%c3%x_d, this%ub%x_d, this%vb%x_d, &
              this%wb%x_d, this%chi%x_d, c%B%x_d, this%K, this%obj_scale, n), &
              'adjoint_forcing')
         ! adv%compute_adjoint (adjoint_pnpn.f90:680-682)
         call this%adv%compute_adjoint(this%u, this%v, this%w, this%ub, this%vb, this%wb, &
              this%fx, this%fy, this%fz)
      end if
      ! explicit source terms (times B) into the forcing; the implicit sink
      ! (source_term%compute, adjoint_pnpn.f90:669)
      call this%source_term%compute(t, this%tstep)
      sink = this%source_term%implicit_sink(t)
      ! (3-4) makeabf + makebdf + lag update
      if (this%time_order .eq. 3) then
         mk3 = [ext(1), ext(2), ext(3), rho / this%dt * bd(2), rho / this%dt * bd(3), &
              rho / this%dt * bd(4)]
         call nk_check(nekoamd_pnpn_makef3(this%Fbx%x_d, this%fxlag%x_d, this%fxlag2%x_d, &
              this%ulag%x_d, this%ulag2%x_d, this%fx%x_d, this%u%x_d, c%B%x_d, mk3, n), &
              'makef3')
         call nk_check(nekoamd_pnpn_makef3(this%Fby%x_d, this%fylag%x_d, this%fylag2%x_d, &
              this%vlag%x_d, this%vlag2%x_d, this%fy%x_d, this%v%x_d, c%B%x_d, mk3, n), &
              'makef3')
         call nk_check(nekoamd_pnpn_makef3(this%Fbz%x_d, this%fzlag%x_d, this%fzlag2%x_d, &
              this%wlag%x_d, this%wlag2%x_d, this%fz%x_d, this%w%x_d, c%B%x_d, mk3, n), &
              'makef3')
      else
         mk = [ext(1), ext(2), rho / this%dt * bd(2), rho / this%dt * bd(3)]
         call nk_check(nekoamd_pnpn_makef_vector(this%Fbx%x_d, this%Fby%x_d, &
              this%Fbz%x_d, this%fxlag%x_d, this%fylag%x_d, this%fzlag%x_d, &
              this%ulag%x_d, this%vlag%x_d, this%wlag%x_d, this%fx%x_d, this%fy%x_d, &
              this%fz%x_d, this%u%x_d, this%v%x_d, this%w%x_d, c%B%x_d, mk, n), &
              'makef_vector')
      end if
      ! (5) pressure residual
      call nk_check(nekoamd_curl(this%t1%x_d, this%t2%x_d, this%t3%x_d, this%ue%x_d, &
           this%ve%x_d, this%we%x_d, c%x%x_d, c%y%x_d, c%z%x_d, Xh%dx_d, lx, nelv), 'curl')
      call mass_average(this, this%t1)
      call mass_average(this, this%t2)
      call mass_average(this, this%t3)
      call nk_check(nekoamd_curl(this%c1%x_d, this%c2%x_d, this%c3%x_d, this%t1%x_d, &
           this%t2%x_d, this%t3%x_d, c%x%x_d, c%y%x_d, c%z%x_d, Xh%dx_d, lx, nelv), 'curl')
      call mass_average(this, this%c1)
      call mass_average(this, this%c2)
      call mass_average(this, this%c3)
      call nk_check(nekoamd_pnpn_prs_force_vector(this%t1%x_d, this%t2%x_d, &
           this%t3%x_d, this%Fbx%x_d, this%Fby%x_d, this%Fbz%x_d, this%c1%x_d, &
           this%c2%x_d, this%c3%x_d, this%ue%x_d, this%ve%x_d, this%we%x_d, sink, &
           c%B%x_d, mu, rho, n), 'prs_force_vector')
      call this%gs_h%op(this%t1, GS_OP_ADD)
      call this%gs_h%op(this%t2, GS_OP_ADD)
      call this%gs_h%op(this%t3, GS_OP_ADD)
      call c%h1%fill(1.0_rp / rho)
      c%ifh2 = .false.
      call this%Ax%compute(this%c1, this%p, c, this%msh, Xh)
      call nk_check(nekoamd_surface_flux(this%c2%x_d, this%u%x_d, this%v%x_d, this%w%x_d, &
           c%x%x_d, c%y%x_d, c%z%x_d, Xh%dx_d, Xh%wx_d, this%face_bits_d, lx, nelv), &
           'surface_flux')
      s = rho * bd(1) / this%dt
      ! p_res = cdtp(Binv t) - A p - s (surface flux): col2, cdtp and sub2s
      ! in one pass
      call nk_check(nekoamd_pnpn_prs_div(this%p_res%x_d, this%t1%x_d, this%t2%x_d, &
           this%t3%x_d, this%Binv%x_d, this%c1%x_d, this%c2%x_d, s, c%x%x_d, c%y%x_d, &
           c%z%x_d, Xh%dx_d, Xh%wx_d, lx, nelv), 'prs_div')
      call this%gs_h%op(this%p_res, GS_OP_ADD)
      call this%bclst_prs%apply_scalar(this%p_res%x_d, n)
      ! (6) pressure solve
      if (.not. this%pc_prs_ready) then
         call this%pc_prs%update()
         this%pc_prs_ready = .true.
      end if
      call this%proj_prs%pre_solving(this%p_res, this%tstep, c, n)
      call this%learned_prs%pre(this%p_res, c, this%bclst_prs, this%gs_h)
      this%gamma_prs = this%learned_prs%gamma
      select type (k => this%ksp_prs)
      type is (fcg_hip_t)
         ! the learned coarse correction inside the flexible CG
         k%mlp = c_null_ptr
         if (this%learned_prs%correct_every > 0 .and. c_associated(this%learned_prs%h)) then
            k%mlp = this%learned_prs%h
            k%every = int(this%learned_prs%correct_every, c_int)
         end if
      end select
      mp = this%ksp_prs%solve(this%Ax, this%dp, this%p_res, n, c, this%bclst_prs, this%gs_h)
      call this%learned_prs%post(this%dp)
      call this%proj_prs%post_solving(this%dp, this%Ax, c, this%bclst_prs, this%gs_h, n, &
           this%tstep)
      call nk_check(nekoamd_add2(this%p%x_d, this%dp%x_d, n), 'add2')
      ! (7) velocity residual, H = mu A + (rho bd0/dt + chi) B
      call c%h1%fill(mu)
      ! h2 = rho bd0/dt + sink in one pass (fill, then add2)
      call nk_check(nekoamd_cadd2(c%h2%x_d, sink, rho * bd(1) / this%dt, n), 'cadd2')
      c%ifh2 = .true.
      call nk_check(nekoamd_opgrad(this%t1%x_d, this%t2%x_d, this%t3%x_d, this%p%x_d, &
           c%x%x_d, c%y%x_d, c%z%x_d, Xh%dx_d, Xh%wx_d, lx, nelv), 'opgrad')
      select type (A => this%Ax)
      type is (ax_helm_hip_t)
         ! the operator on u, v, w and the residual combination in one launch
         call nk_check(nekoamd_pnpn_vel_res_vector(this%c1%x_d, this%c2%x_d, this%c3%x_d, &
              this%u%x_d, this%v%x_d, this%w%x_d, this%Fbx%x_d, this%Fby%x_d, &
              this%Fbz%x_d, this%t1%x_d, this%t2%x_d, this%t3%x_d, Xh%dx_d, c%h1%x_d, &
              c%h2%x_d, c%G11%x_d, c%G12%x_d, c%G13%x_d, c%G22%x_d, c%G23%x_d, &
              c%G33%x_d, c%B%x_d, lx, nelv, 1_c_int), 'vel_res_vector')
      class default
         call this%Ax%compute_vector(this%c1, this%c2, this%c3, this%u, this%v, this%w, &
              c, this%msh, Xh)
         call nk_check(nekoamd_pnpn_vel_res(this%c1%x_d, this%Fbx%x_d, this%t1%x_d, n), &
              'vel_res')
         call nk_check(nekoamd_pnpn_vel_res(this%c2%x_d, this%Fby%x_d, this%t2%x_d, n), &
              'vel_res')
         call nk_check(nekoamd_pnpn_vel_res(this%c3%x_d, this%Fbz%x_d, this%t3%x_d, n), &
              'vel_res')
      end select
      call this%gs_h%op(this%c1, GS_OP_ADD)
      call this%gs_h%op(this%c2, GS_OP_ADD)
      call this%gs_h%op(this%c3, GS_OP_ADD)
      call this%bclst_vel%apply_scalar(this%c1%x_d, n)
      call this%bclst_vel%apply_scalar(this%c2%x_d, n)
      call this%bclst_vel%apply_scalar(this%c3%x_d, n)
      ! (8) velocity solves
      if (bd(1) .ne. this%bd_vel) then
         call this%pc_vel%update()
         this%bd_vel = bd(1)
      end if
      call this%proj_u%pre_solving(this%c1, this%tstep, c, n)
      call this%proj_v%pre_solving(this%c2, this%tstep, c, n)
      call this%proj_w%pre_solving(this%c3, this%tstep, c, n)
      mv = this%ksp_vel%solve_coupled(this%Ax, this%du, this%dv, this%dw, this%c1, &
           this%c2, this%c3, n, c, this%bclst_vel, this%bclst_vel, this%bclst_vel, this%gs_h)
      call this%proj_u%post_solving(this%du, this%Ax, c, this%bclst_vel, this%gs_h, n, &
           this%tstep)
      call this%proj_v%post_solving(this%dv, this%Ax, c, this%bclst_vel, this%gs_h, n, &
           this%tstep)
      call this%proj_w%post_solving(this%dw, this%Ax, c, this%bclst_vel, this%gs_h, n, &
           this%tstep)
      call nk_check(nekoamd_add2(this%u%x_d, this%du%x_d, n), 'add2')
      call nk_check(nekoamd_add2(this%v%x_d, this%dv%x_d, n), 'add2')
      call nk_check(nekoamd_add2(this%w%x_d, this%dw%x_d, n), 'add2')
    end associate
    this%iters = [mp%iter, mv(1)%iter, mv(2)%iter, mv(3)%iter]
    this%res = [mp%res_final, mv(1)%res_final, mv(2)%res_final, mv(3)%res_final]
  end subroutine fluid_step

  subroutine fluid_free(this)
    class(fluid_pnpn_t), intent(inout) :: this
    call this%learned_prs%free()
    call this%brinkman%free()
    call this%source_term%free()
    call this%ub%free()
    call this%vb%free()
    call this%wb%free()
    call this%adv%free()
    this%adjoint = .false.
    call this%proj_prs%free()
    call this%proj_u%free()
    call this%proj_v%free()
    call this%proj_w%free()
    call free_prs_solver(this)
    select type (k => this%ksp_vel)
    type is (cg_hip_t)
       call k%free()
    end select
    select type (p => this%pc_vel)
    type is (jacobi_hip_t)
       call p%free()
    end select
    call this%u%free()
    call this%v%free()
    call this%w%free()
    call this%p%free()
    call this%ulag%free()
    call this%vlag%free()
    call this%wlag%free()
    call this%fxlag%free()
    call this%fylag%free()
    call this%fzlag%free()
    call this%ulag2%free()
    call this%vlag2%free()
    call this%wlag2%free()
    call this%fxlag2%free()
    call this%fylag2%free()
    call this%fzlag2%free()
    call this%ue%free()
    call this%ve%free()
    call this%we%free()
    call this%fx%free()
    call this%fy%free()
    call this%fz%free()
    call this%Fbx%free()
    call this%Fby%free()
    call this%Fbz%free()
    call this%t1%free()
    call this%t2%free()
    call this%t3%free()
    call this%c1%free()
    call this%c2%free()
    call this%c3%free()
    call this%p_res%free()
    call this%dp%free()
    call this%du%free()
    call this%dv%free()
    call this%dw%free()
    call this%chi%free()
    call this%Binv%free()
    call this%h1v%free()
    call this%h2v%free()
    call nk_free(this%face_bits_d)
    call this%bclst_vel%free()
    call this%bclst_prs%free()
    call this%bclst_walls%free()
    call this%c_Xh%free()
    call this%gs_h%free()
    call this%Xh%free()
  end subroutine fluid_free

end module nekoamd_fluid
