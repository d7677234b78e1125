!> The density-based topology-optimisation loop of the reference on the
!! device: one design iteration of examples/rugby_ball/driver.f90:69-138,
!! with the problem's compute / compute_sensitivity of
!! sources/problems/steady_state_problem.f90:273-353:
!!   forward Pn/Pn solve (nfwd steps from the initial state)
!!   objective (minimum dissipation) and volume constraint
!!   adjoint Pn/Pn solve about the forward flow (nadj steps from rest)
!!   dF/dchi, chain rule back through RAMP and the filter (map_backward)
!!   volume sensitivity; both constraint terms scaled by 100 (driver.f90:106-111)
!!   MMA update of the design indicator, KKT residual (driver.f90:113-126)
!!   map_forward of the new design (driver.f90:132)
!! The design indicator starts as seeded noise smoothed by npass-1 filter
!! passes, so that its first map_forward (one more pass, then RAMP) gives
!! the seeded design of the channel case.
module nekoamd_topopt
  use, intrinsic :: iso_c_binding
  use nekoamd_capi
  use nekoamd_types
  use nekoamd_solvers
  use nekoamd_design
  use nekoamd_mma
  use nekoamd_fluid
  use nekoamd_output
  implicit none
  private

  public :: topopt_t

  type :: topopt_t
     type(fluid_pnpn_t) :: fwd, adj
     type(topopt_design_t) :: design
     type(mma_t) :: optimizer
     type(field_t) :: df_dchi, dfdx, xmin, xmax
     real(rp) :: K = 1.0_rp, obj_scale = 1.0_rp, v_max = 0.2_rp
     real(rp) :: volume_domain = 0.0_rp
     logical :: is_max = .false.
     !> last iteration: objective (F, dissipation, lube), volume (constraint,
     !! fraction), MMA (residumax, residunorm, Newton iterations)
     real(rp) :: objective(3) = 0.0_rp, volume(2) = 0.0_rp, kkt(3) = 0.0_rp
     !> pressure iterations summed over the forward and the adjoint solve
     integer :: p_iters(2) = 0
     integer :: iter = 1
     !> the 'optimization' fld series of steady_state_problem.f90:216-244,
     !! sampled every iteration when enabled (rugby_ball/driver.f90:143)
     type(fld_file_output_t) :: output
     logical :: sampling = .false.
   contains
     procedure, pass(this) :: init => topopt_init
     procedure, pass(this) :: iterate => topopt_iterate
     procedure, pass(this) :: free => topopt_free
     procedure, pass(this) :: enable_output => topopt_enable_output
  end type topopt_t

contains

  subroutine topopt_init(this, nx, ny, nz, lx, jitter, seed, re, dt, abstol_p, abstol_v, &
       max_iter, solver, design_seed, npass, f_min, f_max, q, v_max, is_max)
    class(topopt_t), intent(inout), target :: this
    integer, intent(in) :: nx, ny, nz, lx, max_iter, npass
    real(rp), intent(in) :: jitter, re, dt, abstol_p, abstol_v, f_min, f_max, q, v_max
    integer(c_int64_t), intent(in) :: seed, design_seed
    character(len=*), intent(in) :: solver
    logical, intent(in) :: is_max
    integer(c_int64_t) :: n
    integer :: i
    call this%fwd%init(nx, ny, nz, lx, jitter, seed, re, dt, abstol_p, abstol_v, max_iter, &
         solver)
    call this%adj%init(nx, ny, nz, lx, jitter, seed, re, dt, abstol_p, abstol_v, max_iter, &
         solver)
    call this%adj%make_adjoint(this%K, this%obj_scale)
    n = this%fwd%c_Xh%n
    this%v_max = v_max
    this%is_max = is_max
    this%volume_domain = this%fwd%c_Xh%volume
    call this%design%init(this%fwd%c_Xh)
    associate(d_ => this%design)
      call nk_check(nekoamd_box_rng_uniform(d_%design_indicator%x_d, this%fwd%msh%nx, &
           this%fwd%msh%ny, this%fwd%msh%nz, this%fwd%Xh%lx, design_seed), 'rng')
      do i = 1, npass - 1
         call nk_check(nekoamd_copy(d_%sensitivity%x_d, d_%design_indicator%x_d, n), 'copy')
         call d_%filter%apply_forward(d_%design_indicator, d_%sensitivity)
      end do
      d_%mapping%f_min = f_min
      d_%mapping%f_max = f_max
      d_%mapping%q = q
      d_%mapping%convex_up = .true.
    end associate
    call set_design(this)
    call this%df_dchi%init(n, 'df_dchi')
    call this%dfdx%init(n, 'dfdx')
    call this%xmin%init(n, 'xmin')
    call this%xmax%init(n, 'xmax')
    call this%xmax%fill(1.0_rp)
    ! optimizer%init_json(x, n, 1, a0 = 1, a = 0, c = 100, d = 0, xmin = 0,
    ! xmax = 1) with the json defaults (driver.f90:51-53, mma.f90:160-176)
    call this%optimizer%init(this%design%design_indicator, int(n), 1, 1.0_rp, [0.0_rp], &
         [100.0_rp], [0.0_rp], this%xmin, this%xmax)
    this%iter = 1
  end subroutine topopt_init

  !> The 13 fields of the reference's optimisation output, in its order
  !! (steady_state_problem.f90:216-244): p, u, v, w, design indicator
  !! (written as temperature), adjoint u, v, w, p, chi, dF/dchi, the volume
  !! constraint's sensitivity, dF/drho (scalars s1..s8).
  subroutine topopt_enable_output(this, name, wdsize)
    class(topopt_t), intent(inout), target :: this
    character(len=*), intent(in) :: name
    integer, intent(in) :: wdsize
    call this%output%init(this%fwd%c_Xh, name, 13, wdsize)
    call this%output%assign(1, this%fwd%p)
    call this%output%assign(2, this%fwd%u)
    call this%output%assign(3, this%fwd%v)
    call this%output%assign(4, this%fwd%w)
    call this%output%assign(5, this%design%design_indicator)
    call this%output%assign(6, this%adj%u)
    call this%output%assign(7, this%adj%v)
    call this%output%assign(8, this%adj%w)
    call this%output%assign(9, this%adj%p)
    call this%output%assign(10, this%design%brinkman_amplitude)
    call this%output%assign(11, this%df_dchi)
    call this%output%assign(12, this%dfdx)
    call this%output%assign(13, this%design%sensitivity)
    this%sampling = .true.
  end subroutine topopt_enable_output

  !> map_forward and hand chi to both solves
  subroutine set_design(this)
    type(topopt_t), intent(inout) :: this
    integer(c_int64_t) :: n
    n = this%fwd%c_Xh%n
    call this%design%map_forward()
    call nk_check(nekoamd_copy(this%fwd%chi%x_d, this%design%brinkman_amplitude%x_d, n), &
         'copy')
    call nk_check(nekoamd_copy(this%adj%chi%x_d, this%design%brinkman_amplitude%x_d, n), &
         'copy')
    this%fwd%bd_vel = -1.0_rp
    this%adj%bd_vel = -1.0_rp
  end subroutine set_design

  subroutine topopt_iterate(this, nfwd, nadj)
    class(topopt_t), intent(inout) :: this
    integer, intent(in) :: nfwd, nadj
    integer(c_int64_t) :: n
    integer(c_int) :: lx, nelv, lube, ismax
    real(c_double) :: fval(1), s, vd, vm
    integer :: k
    n = this%fwd%c_Xh%n
    lx = this%fwd%c_Xh%lx
    nelv = this%fwd%c_Xh%nelv
    lube = 1
    ismax = merge(1, 0, this%is_max)
    vd = this%volume_domain
    vm = this%v_max
    ! problem%compute: forward solve, objective, volume (steady_state_problem.f90:273-306)
    call this%fwd%reset()
    this%p_iters = 0
    do k = 1, nfwd
       call this%fwd%step()
       this%p_iters(1) = this%p_iters(1) + this%fwd%iters(1)
    end do
    associate(f => this%fwd, c => this%fwd%c_Xh)
      call nk_check(nekoamd_dissipation_objective(this%objective, f%u%x_d, f%v%x_d, f%w%x_d, &
           f%chi%x_d, c%B%x_d, c%x%x_d, c%y%x_d, c%z%x_d, c%Xh%dx_d, lx, nelv, this%K, &
           this%obj_scale, lube), 'objective')
      call nk_check(nekoamd_volume_constraint(this%volume, &
           this%design%design_indicator%x_d, c%B%x_d, vd, vm, ismax, n), 'volume')
    end associate
    ! problem%compute_sensitivity: adjoint solve about the forward flow (:309-353)
    call this%adj%reset()
    call nk_check(nekoamd_copy(this%adj%ub%x_d, this%fwd%u%x_d, n), 'copy')
    call nk_check(nekoamd_copy(this%adj%vb%x_d, this%fwd%v%x_d, n), 'copy')
    call nk_check(nekoamd_copy(this%adj%wb%x_d, this%fwd%w%x_d, n), 'copy')
    do k = 1, nadj
       call this%adj%step()
       this%p_iters(2) = this%p_iters(2) + this%adj%iters(1)
    end do
    call nk_check(nekoamd_dissipation_sensitivity(this%df_dchi%x_d, this%fwd%u%x_d, &
         this%fwd%v%x_d, this%fwd%w%x_d, this%adj%u%x_d, this%adj%v%x_d, this%adj%w%x_d, &
         this%K, this%obj_scale, lube, n), 'sensitivity')
    call this%design%map_backward(this%df_dchi)
    call nk_check(nekoamd_volume_sensitivity(this%dfdx%x_d, vd, ismax, n), &
         'volume_sensitivity')
    ! scaled constraint (driver.f90:106-111)
    s = 100.0_rp
    call nk_check(nekoamd_cmult(this%dfdx%x_d, s, n), 'cmult')
    fval(1) = this%volume(1) * 100.0_rp
    if (this%sampling) call this%output%sample(real(this%iter, rp))
    ! optimizer%mma_update, kkt (driver.f90:113-126)
    call this%optimizer%update(this%iter, this%design%design_indicator, &
         this%design%sensitivity, fval, this%dfdx)
    call this%optimizer%KKT(this%design%design_indicator, this%design%sensitivity, fval, &
         this%dfdx)
    this%kkt = [this%optimizer%get_residumax(), this%optimizer%get_residunorm(), &
         real(this%optimizer%get_newton_iters(), rp)]
    this%iter = this%iter + 1
    call set_design(this)
  end subroutine topopt_iterate

  subroutine topopt_free(this)
    class(topopt_t), intent(inout) :: this
    call this%optimizer%free()
    call this%output%free()
    this%sampling = .false.
    call this%df_dchi%free()
    call this%dfdx%free()
    call this%xmin%free()
    call this%xmax%free()
    call this%design%free()
    call this%adj%free()
    call this%fwd%free()
  end subroutine topopt_free

end module nekoamd_topopt
