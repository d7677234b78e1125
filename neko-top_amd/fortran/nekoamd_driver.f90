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
  use nekoamd_design
  use nekoamd_fluid
  use nekoamd_topopt
  use nekoamd_output
  implicit none
  private

  logical, save :: device_ready = .false.

  !> Channel cases driven through the C-ABI: handle = slot index.
  type :: channel_slot
     type(fluid_pnpn_t), pointer :: f => null()
     type(topopt_design_t), pointer :: d => null()
     type(fld_file_output_t), pointer :: out => null()
  end type channel_slot
  integer, parameter :: max_slots = 8
  type(channel_slot), save :: slots(max_slots)

  !> Optimisation loops driven through the C-ABI: handle = slot index.
  type :: topopt_slot
     type(topopt_t), pointer :: t => null()
  end type topopt_slot
  integer, parameter :: max_topopt = 2
  type(topopt_slot), save :: tslots(max_topopt)

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
    type(box_mesh_t), target :: msh
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


  function slot_of(handle) result(k)
    integer(c_int), intent(in) :: handle
    integer :: k
    k = handle
    if (k .lt. 1 .or. k .gt. max_slots) error stop 'channel: invalid handle'
    if (.not. associated(slots(k)%f)) error stop 'channel: handle not in use'
  end function slot_of

  !> Brinkman channel on an nx*ny*nz box of order lx-1 (uniform inflow at
  !! x = 0, outflow at x = L, no-slip walls), chi = 0 until a design is set.
  !! precision: NK_CG_FP64 ('cg') or NK_CG_MIXED ('cg_mixed') solves.
  function nekoamd_f_channel_create(nx, ny, nz, lx, jitter, seed, re, dt, abstol_p, &
       abstol_v, max_iter, precision, handle) result(rc) &
       bind(c, name='nekoamd_f_channel_create')
    integer(c_int) :: nx, ny, nz, lx, max_iter, precision, handle
    real(c_double) :: jitter, re, dt, abstol_p, abstol_v
    integer(c_int64_t) :: seed
    integer(c_int) :: rc
    integer :: k
    call ensure_device()
    handle = 0
    rc = 1
    if (precision .ne. NK_CG_FP64 .and. precision .ne. NK_CG_MIXED) return
    do k = 1, max_slots
       if (.not. associated(slots(k)%f)) then
          allocate(slots(k)%f)
          call slots(k)%f%init(int(nx), int(ny), int(nz), int(lx), jitter, seed, re, dt, &
               abstol_p, abstol_v, int(max_iter), &
               merge('cg_mixed', 'cg      ', precision .eq. NK_CG_MIXED))
          handle = k
          rc = 0
          return
       end if
    end do
  end function nekoamd_f_channel_create

  !> Projection spaces of the channel's solves: L_prs / L_vel previous
  !! solutions (0 = off), active after activ_step steps; keep_prs / keep_vel
  !! solutions kept when a full basis restarts (0: Neko's restart).
  function nekoamd_f_channel_projection(handle, L_prs, L_vel, activ_step, keep_prs, &
       keep_vel) result(rc) bind(c, name='nekoamd_f_channel_projection')
    integer(c_int) :: handle, L_prs, L_vel, activ_step, keep_prs, keep_vel
    integer(c_int) :: rc
    integer :: k
    k = slot_of(handle)
    rc = 1
    if (L_prs .lt. 0 .or. L_vel .lt. 0 .or. L_prs .gt. 64 .or. L_vel .gt. 64) return
    if (keep_prs .lt. 0 .or. keep_vel .lt. 0) return
    if (keep_prs .gt. 0 .and. keep_prs .ge. L_prs) return
    if (keep_vel .gt. 0 .and. keep_vel .ge. L_vel) return
    call slots(k)%f%projection(int(L_prs), int(L_vel), int(activ_step), &
         int(keep_prs), int(keep_vel))
    rc = 0
  end function nekoamd_f_channel_projection

  !> The projection bases in FP32 (on 1) or FP64 (on 0), after
  !! nekoamd_f_channel_projection and before the first step.
  function nekoamd_f_channel_projection_fp32(handle, on) result(rc) &
       bind(c, name='nekoamd_f_channel_projection_fp32')
    integer(c_int) :: handle, on
    integer(c_int) :: rc
    integer :: k
    k = slot_of(handle)
    associate(f => slots(k)%f)
      call f%proj_prs%set_fp32(on .ne. 0)
      call f%proj_u%set_fp32(on .ne. 0)
      call f%proj_v%set_fp32(on .ne. 0)
      call f%proj_w%set_fp32(on .ne. 0)
    end associate
    rc = 0
  end function nekoamd_f_channel_projection_fp32

  !> Learned guess of the pressure solves: mode 0 off, 1 harvest, 2 apply,
  !! 3 apply and harvest.  The network is created on the first call with
  !! capacity harvested pairs, batch training rows and the weight seed.
  function nekoamd_f_channel_learned(handle, mode, capacity, batch, seed) result(rc) &
       bind(c, name='nekoamd_f_channel_learned')
    integer(c_int) :: handle, mode, capacity, batch
    integer(c_int64_t) :: seed
    integer(c_int) :: rc
    integer :: k
    k = slot_of(handle)
    rc = 1
    if (mode .lt. LEARNED_OFF .or. mode .gt. LEARNED_APPLY_HARVEST) return
    associate(f => slots(k)%f)
      if (.not. c_associated(f%learned_prs%h)) &
           call f%learned_prs%init(f%msh%nx, f%msh%ny, f%msh%nz, f%Xh%lx, int(capacity), &
           int(batch), seed)
      f%learned_prs%mode = mode
    end associate
    rc = 0
  end function nekoamd_f_channel_learned

  !> Train the pressure network on its harvested pairs; mean loss of the
  !! first and last epoch; npairs = pairs trained on.
  function nekoamd_f_channel_train(handle, epochs, batch, lr, seed, loss0, loss1, npairs) &
       result(rc) bind(c, name='nekoamd_f_channel_train')
    integer(c_int) :: handle, epochs, batch, npairs
    real(c_float) :: lr
    integer(c_int64_t) :: seed
    real(c_double) :: loss0, loss1
    integer(c_int) :: rc
    integer :: k
    k = slot_of(handle)
    npairs = slots(k)%f%learned_prs%pairs()
    call slots(k)%f%learned_prs%train(int(epochs), int(batch), lr, seed, loss0, loss1)
    rc = 0
  end function nekoamd_f_channel_train

  !> The network's correction inside the flexible-CG pressure solve every
  !! `every`-th iteration (0: off); needs set_pressure_solver('fcg', 'jacobi')
  !! and a network (nekoamd_f_channel_learned).
  function nekoamd_f_channel_learned_correction(handle, every) result(rc) &
       bind(c, name='nekoamd_f_channel_learned_correction')
    integer(c_int) :: handle, every
    integer(c_int) :: rc
    integer :: k
    k = slot_of(handle)
    if (every > 0 .and. .not. c_associated(slots(k)%f%learned_prs%h)) &
         error stop 'channel: the learned correction needs a network (learned)'
    slots(k)%f%learned_prs%correct_every = max(0, int(every))
    rc = 0
  end function nekoamd_f_channel_learned_correction

  !> gamma of the last pressure solve's learned direction (0: none applied).
  function nekoamd_f_channel_gamma(handle, gamma) result(rc) &
       bind(c, name='nekoamd_f_channel_gamma')
    integer(c_int) :: handle
    real(c_double) :: gamma
    integer(c_int) :: rc
    integer :: k
    k = slot_of(handle)
    gamma = slots(k)%f%gamma_prs
    rc = 0
  end function nekoamd_f_channel_gamma

  function topopt_of(handle) result(k)
    integer(c_int), intent(in) :: handle
    integer :: k
    k = handle
    if (k .lt. 1 .or. k .gt. max_topopt) error stop 'topopt: invalid handle'
    if (.not. associated(tslots(k)%t)) error stop 'topopt: handle not in use'
  end function topopt_of

  !> The optimisation loop on a channel case (forward and adjoint Pn/Pn
  !! solves of precision NK_CG_FP64 or NK_CG_MIXED, design seeded as the
  !! channel's, volume constraint v_max / is_max).
  function nekoamd_f_topopt_create(nx, ny, nz, lx, jitter, seed, re, dt, abstol_p, &
       abstol_v, max_iter, precision, design_seed, npass, f_min, f_max, q, v_max, is_max, &
       handle) result(rc) bind(c, name='nekoamd_f_topopt_create')
    integer(c_int) :: nx, ny, nz, lx, max_iter, precision, npass, is_max, handle
    real(c_double) :: jitter, re, dt, abstol_p, abstol_v, f_min, f_max, q, v_max
    integer(c_int64_t) :: seed, design_seed
    integer(c_int) :: rc
    integer :: k
    call ensure_device()
    handle = 0
    rc = 1
    if (precision .ne. NK_CG_FP64 .and. precision .ne. NK_CG_MIXED) return
    if (npass .lt. 1) return
    do k = 1, max_topopt
       if (.not. associated(tslots(k)%t)) then
          allocate(tslots(k)%t)
          call tslots(k)%t%init(int(nx), int(ny), int(nz), int(lx), jitter, seed, re, dt, &
               abstol_p, abstol_v, int(max_iter), &
               merge('cg_mixed', 'cg      ', precision .eq. NK_CG_MIXED), design_seed, &
               int(npass), f_min, f_max, q, v_max, is_max .ne. 0)
          handle = k
          rc = 0
          return
       end if
    end do
  end function nekoamd_f_topopt_create

  !> One design iteration (nfwd forward and nadj adjoint steps).
  !! out: objective F, dissipation, lube, volume constraint, volume fraction,
  !! KKT residumax, residunorm, MMA Newton iterations (of the new design's
  !! update); iters: pressure iterations of the forward and adjoint solves.
  function nekoamd_f_topopt_iterate(handle, nfwd, nadj, out, iters) result(rc) &
       bind(c, name='nekoamd_f_topopt_iterate')
    integer(c_int) :: handle, nfwd, nadj
    real(c_double) :: out(8)
    integer(c_int) :: iters(2)
    integer(c_int) :: rc
    integer :: k
    k = topopt_of(handle)
    associate(t => tslots(k)%t)
      call t%iterate(int(nfwd), int(nadj))
      out(1:3) = t%objective
      out(4:5) = t%volume
      out(6:8) = t%kkt
      iters = t%p_iters
    end associate
    rc = 0
  end function nekoamd_f_topopt_iterate

  !> which: 0 design indicator, 1 filtered design, 2 chi, 3 dF/drho,
  !! 4 dF/dchi, 5-7 forward u, v, w, 8-10 adjoint u, v, w
  function nekoamd_f_topopt_get(handle, which, out) result(rc) &
       bind(c, name='nekoamd_f_topopt_get')
    integer(c_int) :: handle, which
    real(c_double), target :: out(*)
    integer(c_int) :: rc
    integer :: k
    type(c_ptr) :: p
    k = topopt_of(handle)
    associate(t => tslots(k)%t)
      select case (which)
      case (0); p = t%design%design_indicator%x_d
      case (1); p = t%design%filtered_design%x_d
      case (2); p = t%design%brinkman_amplitude%x_d
      case (3); p = t%design%sensitivity%x_d
      case (4); p = t%df_dchi%x_d
      case (5); p = t%fwd%u%x_d
      case (6); p = t%fwd%v%x_d
      case (7); p = t%fwd%w%x_d
      case (8); p = t%adj%u%x_d
      case (9); p = t%adj%v%x_d
      case (10); p = t%adj%w%x_d
      case default
         error stop 'topopt: unknown field'
      end select
      call nk_download(c_loc(out), p, 8_c_int64_t * t%fwd%c_Xh%n)
    end associate
    rc = 0
  end function nekoamd_f_topopt_get

  function nekoamd_f_topopt_free(handle) result(rc) bind(c, name='nekoamd_f_topopt_free')
    integer(c_int) :: handle
    integer(c_int) :: rc
    integer :: k
    k = topopt_of(handle)
    call tslots(k)%t%free()
    deallocate(tslots(k)%t)
    rc = 0
  end function nekoamd_f_topopt_free

  !> Sample p, u, v, w of the channel into the fld series `name` (a C
  !! string; the first call, or a call with another name or wdsize, starts a
  !! new series), at time t.
  function nekoamd_f_channel_sample(handle, name, t, wdsize) result(rc) &
       bind(c, name='nekoamd_f_channel_sample')
    integer(c_int) :: handle, wdsize
    character(kind=c_char), dimension(*) :: name
    real(c_double) :: t
    integer(c_int) :: rc
    integer :: k, l
    character(len=:), allocatable :: nm
    k = slot_of(handle)
    l = 0
    do while (name(l + 1) .ne. c_null_char)
       l = l + 1
    end do
    allocate(character(len=l) :: nm)
    nm = transfer(name(1:l), nm)
    if (.not. associated(slots(k)%out)) allocate(slots(k)%out)
    associate(o => slots(k)%out, f => slots(k)%f)
      if (.not. allocated(o%name)) then
         call start()
      else if (o%name .ne. nm .or. o%wdsize .ne. wdsize) then
         call start()
      end if
      call o%sample(t)
    end associate
    rc = 0
  contains
    subroutine start()
      associate(o => slots(k)%out, f => slots(k)%f)
        call o%init(f%c_Xh, nm, 4, int(wdsize))
        call o%assign(1, f%p)
        call o%assign(2, f%u)
        call o%assign(3, f%v)
        call o%assign(4, f%w)
      end associate
    end subroutine start
  end function nekoamd_f_channel_sample

  !> The pressure solver of the channel: solver 'cg' | 'cg_mixed' | 'gmres',
  !! pc 'jacobi' | 'hsmg' (C strings), e.g. gmres + hsmg as the shipped cases
  !! run it (rugby_ball.case:46-52).
  function nekoamd_f_channel_pressure_solver(handle, solver, pc) result(rc) &
       bind(c, name='nekoamd_f_channel_pressure_solver')
    integer(c_int) :: handle
    character(kind=c_char), dimension(*) :: solver, pc
    integer(c_int) :: rc
    integer :: k
    k = slot_of(handle)
    call slots(k)%f%set_pressure_solver(c_to_f(solver), c_to_f(pc))
    rc = 0
  end function nekoamd_f_channel_pressure_solver

  !> both solves of the optimisation loop
  function nekoamd_f_topopt_pressure_solver(handle, solver, pc) result(rc) &
       bind(c, name='nekoamd_f_topopt_pressure_solver')
    integer(c_int) :: handle
    character(kind=c_char), dimension(*) :: solver, pc
    integer(c_int) :: rc
    integer :: k
    k = topopt_of(handle)
    call tslots(k)%t%fwd%set_pressure_solver(c_to_f(solver), c_to_f(pc))
    call tslots(k)%t%adj%set_pressure_solver(c_to_f(solver), c_to_f(pc))
    rc = 0
  end function nekoamd_f_topopt_pressure_solver

  function c_to_f(s) result(f)
    character(kind=c_char), dimension(*), intent(in) :: s
    character(len=:), allocatable :: f
    integer :: l, i
    l = 0
    do while (s(l + 1) .ne. c_null_char)
       l = l + 1
    end do
    allocate(character(len=l) :: f)
    do i = 1, l
       f(i:i) = s(i)
    end do
  end function c_to_f

  !> The optimisation loop's 13-field 'optimization' series, sampled every
  !! design iteration (steady_state_problem.f90:216-244).
  function nekoamd_f_topopt_output(handle, name, wdsize) result(rc) &
       bind(c, name='nekoamd_f_topopt_output')
    integer(c_int) :: handle, wdsize
    character(kind=c_char), dimension(*) :: name
    integer(c_int) :: rc
    integer :: k, l
    character(len=:), allocatable :: nm
    k = topopt_of(handle)
    l = 0
    do while (name(l + 1) .ne. c_null_char)
       l = l + 1
    end do
    allocate(character(len=l) :: nm)
    nm = transfer(name(1:l), nm)
    call tslots(k)%t%enable_output(nm, int(wdsize))
    rc = 0
  end function nekoamd_f_topopt_output

  !> The Brinkman term of the channel: implicit = 1 semi-implicit (chi in the
  !! velocity Helmholtz operator), 0 the reference's explicit source term.
  function nekoamd_f_channel_brinkman(handle, implicit) result(rc) &
       bind(c, name='nekoamd_f_channel_brinkman')
    integer(c_int) :: handle, implicit
    integer(c_int) :: rc
    integer :: k
    k = slot_of(handle)
    call slots(k)%f%set_brinkman(implicit .ne. 0)
    rc = 0
  end function nekoamd_f_channel_brinkman

  !> case.numerics.dealias for the channel: on = 1 puts the forward and
  !! adjoint advection terms on lxd Gauss-Legendre points (lxd <= 0: the
  !! reference's default 3 (p + 1) / 2), on = 0 back to the collocated forms.
  function nekoamd_f_channel_dealias(handle, on, lxd) result(rc) &
       bind(c, name='nekoamd_f_channel_dealias')
    integer(c_int) :: handle, on, lxd
    integer(c_int) :: rc
    integer :: k
    k = slot_of(handle)
    call slots(k)%f%set_dealias(on .ne. 0, int(lxd))
    rc = 0
  end function nekoamd_f_channel_dealias

  !> The same for both solves of an optimisation loop.
  function nekoamd_f_topopt_dealias(handle, on, lxd) result(rc) &
       bind(c, name='nekoamd_f_topopt_dealias')
    integer(c_int) :: handle, on, lxd
    integer(c_int) :: rc
    integer :: k
    k = topopt_of(handle)
    call tslots(k)%t%fwd%set_dealias(on .ne. 0, int(lxd))
    call tslots(k)%t%adj%set_dealias(on .ne. 0, int(lxd))
    rc = 0
  end function nekoamd_f_topopt_dealias

  !> Turn the channel into the adjoint solve (adjoint_pnpn_step): state
  !! zeroed; the base flow (fields 7-9) is set with nekoamd_f_channel_set.
  function nekoamd_f_channel_adjoint(handle, kcoef, obj_scale) result(rc) &
       bind(c, name='nekoamd_f_channel_adjoint')
    integer(c_int) :: handle
    real(c_double) :: kcoef, obj_scale
    integer(c_int) :: rc
    integer :: k
    k = slot_of(handle)
    call slots(k)%f%make_adjoint(kcoef, obj_scale)
    rc = 0
  end function nekoamd_f_channel_adjoint

  !> Chain rule of the design maps (topopt_design.f90:316-352): given
  !! df_dchi (host), sens = filter^T(RAMP'(rho~) df_dchi) (host).  Needs a
  !! design (nekoamd_f_channel_design).
  function nekoamd_f_channel_map_backward(handle, df_dchi, sens) result(rc) &
       bind(c, name='nekoamd_f_channel_map_backward')
    integer(c_int) :: handle
    real(c_double), target :: df_dchi(*), sens(*)
    integer(c_int) :: rc
    integer :: k
    type(field_t) :: g
    k = slot_of(handle)
    if (.not. associated(slots(k)%d)) error stop 'channel: no design'
    call g%init(slots(k)%f%c_Xh%n, 'df_dchi')
    call g%upload(df_dchi)
    call slots(k)%d%map_backward(g)
    call nk_download(c_loc(sens), slots(k)%d%sensitivity%x_d, &
         8_c_int64_t * slots(k)%f%c_Xh%n)
    call g%free()
    rc = 0
  end function nekoamd_f_channel_map_backward

  !> Seeded design: rho ~ U[0,1) per global point (seed), npass PDE-filter
  !! passes, chi = RAMP(rho~) (convex-up with f_min, f_max, q).
  function nekoamd_f_channel_design(handle, seed, npass, f_min, f_max, q) result(rc) &
       bind(c, name='nekoamd_f_channel_design')
    integer(c_int) :: handle, npass
    integer(c_int64_t) :: seed
    real(c_double) :: f_min, f_max, q
    integer(c_int) :: rc
    integer :: k, i
    type(fluid_pnpn_t), pointer :: f
    type(topopt_design_t), pointer :: d
    k = slot_of(handle)
    f => slots(k)%f
    if (.not. associated(slots(k)%d)) then
       allocate(slots(k)%d)
       call slots(k)%d%init(f%c_Xh)
    end if
    d => slots(k)%d
    call nk_check(nekoamd_box_rng_uniform(d%design_indicator%x_d, f%msh%nx, f%msh%ny, &
         f%msh%nz, f%Xh%lx, seed), 'rng')
    call nk_check(nekoamd_copy(d%filtered_design%x_d, d%design_indicator%x_d, &
         f%c_Xh%n), 'copy')
    do i = 1, npass
       call nk_check(nekoamd_copy(d%sensitivity%x_d, d%filtered_design%x_d, f%c_Xh%n), 'copy')
       call d%filter%apply_forward(d%filtered_design, d%sensitivity)
    end do
    d%mapping%f_min = f_min
    d%mapping%f_max = f_max
    d%mapping%q = q
    d%mapping%convex_up = .true.
    call d%mapping%apply_forward(f%chi, d%filtered_design)
    f%bd_vel = -1.0_rp
    rc = 0
  end function nekoamd_f_channel_design

  !> nsteps time steps; iters(4*s+1..4*s+4) = pressure, u, v, w iterations of
  !! step s, res likewise the final residuals.
  function nekoamd_f_channel_step(handle, nsteps, iters, res) result(rc) &
       bind(c, name='nekoamd_f_channel_step')
    integer(c_int) :: handle, nsteps
    integer(c_int) :: iters(*)
    real(c_double) :: res(*)
    integer(c_int) :: rc
    integer :: k, s
    k = slot_of(handle)
    do s = 0, nsteps - 1
       call slots(k)%f%step()
       iters(4 * s + 1:4 * s + 4) = slots(k)%f%iters
       res(4 * s + 1:4 * s + 4) = slots(k)%f%res
    end do
    rc = 0
  end function nekoamd_f_channel_step

  function channel_field(k, which) result(p)
    integer, intent(in) :: k
    integer(c_int), intent(in) :: which
    type(c_ptr) :: p
    select case (which)
    case (0); p = slots(k)%f%u%x_d
    case (1); p = slots(k)%f%v%x_d
    case (2); p = slots(k)%f%w%x_d
    case (3); p = slots(k)%f%p%x_d
    case (4); p = slots(k)%f%chi%x_d
    case (5)
       if (.not. associated(slots(k)%d)) error stop 'channel: no design'
       p = slots(k)%d%design_indicator%x_d
    case (6)
       if (.not. associated(slots(k)%d)) error stop 'channel: no design'
       p = slots(k)%d%filtered_design%x_d
    case (7)
       if (.not. slots(k)%f%adjoint) error stop 'channel: not an adjoint case'
       p = slots(k)%f%ub%x_d
    case (8)
       if (.not. slots(k)%f%adjoint) error stop 'channel: not an adjoint case'
       p = slots(k)%f%vb%x_d
    case (9)
       if (.not. slots(k)%f%adjoint) error stop 'channel: not an adjoint case'
       p = slots(k)%f%wb%x_d
    case (10); p = slots(k)%f%ulag%x_d
    case (11); p = slots(k)%f%vlag%x_d
    case (12); p = slots(k)%f%wlag%x_d
    case (13); p = slots(k)%f%fxlag%x_d
    case (14); p = slots(k)%f%fylag%x_d
    case (15); p = slots(k)%f%fzlag%x_d
    case (16:21)
       if (slots(k)%f%time_order .ne. 3) error stop 'channel: second lags need time_order 3'
       select case (which)
       case (16); p = slots(k)%f%ulag2%x_d
       case (17); p = slots(k)%f%vlag2%x_d
       case (18); p = slots(k)%f%wlag2%x_d
       case (19); p = slots(k)%f%fxlag2%x_d
       case (20); p = slots(k)%f%fylag2%x_d
       case default; p = slots(k)%f%fzlag2%x_d
       end select
    case default
       error stop 'channel: unknown field'
    end select
  end function channel_field

  !> which: 0 u, 1 v, 2 w, 3 p, 4 chi, 5 rho, 6 filtered rho, 7-9 base flow
  !! ub, vb, wb of an adjoint case, 10-12 the velocity lags, 13-15 the lagged
  !! explicit forcing, 16-21 their second lags (time_order 3)
  function nekoamd_f_channel_get(handle, which, out) result(rc) &
       bind(c, name='nekoamd_f_channel_get')
    integer(c_int) :: handle, which
    real(c_double), target :: out(*)
    integer(c_int) :: rc
    integer :: k
    k = slot_of(handle)
    call nk_download(c_loc(out), channel_field(k, which), 8_c_int64_t * slots(k)%f%c_Xh%n)
    rc = 0
  end function nekoamd_f_channel_get

  function nekoamd_f_channel_set(handle, which, in) result(rc) &
       bind(c, name='nekoamd_f_channel_set')
    integer(c_int) :: handle, which
    real(c_double), target :: in(*)
    integer(c_int) :: rc
    integer :: k
    k = slot_of(handle)
    call nk_upload(channel_field(k, which), c_loc(in), 8_c_int64_t * slots(k)%f%c_Xh%n)
    if (which .eq. 4) slots(k)%f%bd_vel = -1.0_rp
    rc = 0
  end function nekoamd_f_channel_set

  !> case.numerics.time_order of a channel that has not stepped yet: 2
  !! (BDF2/EXT2) or 3 (BDF3/EXT3 from the third step, as the reference's
  !! cases run).
  function nekoamd_f_channel_time_order(handle, order) result(rc) &
       bind(c, name='nekoamd_f_channel_time_order')
    integer(c_int) :: handle, order
    integer(c_int) :: rc
    integer :: k
    k = slot_of(handle)
    call slots(k)%f%set_time_order(int(order))
    rc = 0
  end function nekoamd_f_channel_time_order

  !> The step counter: get (set = 0) or set (set = 1; the next step is
  !! tstep + 1).
  function nekoamd_f_channel_tstep(handle, tstep, set) result(rc) &
       bind(c, name='nekoamd_f_channel_tstep')
    integer(c_int) :: handle, tstep, set
    integer(c_int) :: rc
    integer :: k
    k = slot_of(handle)
    if (set .ne. 0) then
       slots(k)%f%tstep = int(tstep)
    else
       tstep = int(slots(k)%f%tstep, c_int)
    end if
    rc = 0
  end function nekoamd_f_channel_tstep

  !> Checkpoint (restart = 0) or restart (restart = 1) the channel's state
  !! through fluid_pnpn_t%checkpoint / %restart; path is NUL-terminated.
  function nekoamd_f_channel_chkp(handle, path, restart) result(rc) &
       bind(c, name='nekoamd_f_channel_chkp')
    integer(c_int) :: handle, restart
    character(kind=c_char), intent(in) :: path(*)
    integer(c_int) :: rc
    character(len=:), allocatable :: p
    integer :: k, l, i
    k = slot_of(handle)
    l = 0
    do while (path(l + 1) .ne. c_null_char)
       l = l + 1
    end do
    allocate(character(len=l) :: p)
    do i = 1, l
       p(i:i) = path(i)
    end do
    if (restart .ne. 0) then
       call slots(k)%f%restart(p)
    else
       call slots(k)%f%checkpoint(p)
    end if
    rc = 0
  end function nekoamd_f_channel_chkp

  function nekoamd_f_channel_free(handle) result(rc) bind(c, name='nekoamd_f_channel_free')
    integer(c_int) :: handle
    integer(c_int) :: rc
    integer :: k
    k = slot_of(handle)
    if (associated(slots(k)%d)) then
       call slots(k)%d%free()
       deallocate(slots(k)%d)
    end if
    if (associated(slots(k)%out)) then
       call slots(k)%out%free()
       deallocate(slots(k)%out)
    end if
    call slots(k)%f%free()
    deallocate(slots(k)%f)
    rc = 0
  end function nekoamd_f_channel_free

end module nekoamd_driver
