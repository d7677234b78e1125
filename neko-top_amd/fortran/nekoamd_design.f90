!> Design maps of the topology optimisation, on the device:
!!   pde_filter_t   rho -> rho~ by -r^2 lap(rho~) + rho~ = rho, Neumann
!!                  (sources/mapping_functions/PDE_filter_mapping.f90:72-363)
!!   ramp_mapping_t rho~ -> chi and its chain rule
!!                  (sources/mapping_functions/RAMP_mapping.f90:72-267)
!!   topopt_design_t map_forward / map_backward
!!                  (sources/designs/topopt_design.f90:292-352)
!! The filter solves with the same ax_helm + PCG as the flow (the reference
!! uses GMRES with the identity preconditioner; the solution of the SPD
!! system is the same to the tolerance) and, as the reference does, sets the
!! Helmholtz coefficients of the shared coef_t (h1 = r^2, h2 = 1) before its
!! solve.  Its right-hand side is assembled (gs) and it has no Dirichlet
!! points.
module nekoamd_design
  use, intrinsic :: iso_c_binding
  use nekoamd_capi
  use nekoamd_types
  use nekoamd_solvers
  implicit none
  private

  public :: mapping_t, pde_filter_t, ramp_mapping_t, topopt_design_t

  !> The reference's abstract mapping (sources/mapping_functions/
  !! mapping.f90:44-61): rho -> rho~ forward, and the chain rule backward
  !! dF/drho~ -> dF/drho.  init takes the coefficients only (the reference
  !! also reads a JSON dictionary; the parameters here are components set
  !! before init).
  type, abstract :: mapping_t
     type(coef_t), pointer :: coef => null()
   contains
     procedure, pass(this) :: init_base => mapping_init_base
     procedure, pass(this) :: free_base => mapping_free_base
     procedure(mapping_init_i), pass(this), deferred :: init
     procedure(mapping_free_i), pass(this), deferred :: free
     procedure(mapping_apply_i), pass(this), deferred :: apply_forward
     procedure(mapping_apply_backward_i), pass(this), deferred :: apply_backward
  end type mapping_t

  abstract interface
     subroutine mapping_init_i(this, coef)
       import mapping_t, coef_t
       class(mapping_t), intent(inout) :: this
       type(coef_t), intent(inout), target :: coef
     end subroutine mapping_init_i
     subroutine mapping_free_i(this)
       import mapping_t
       class(mapping_t), intent(inout) :: this
     end subroutine mapping_free_i
     subroutine mapping_apply_i(this, X_out, X_in)
       import mapping_t, field_t
       class(mapping_t), intent(inout) :: this
       type(field_t), intent(inout) :: X_out
       type(field_t), intent(in) :: X_in
     end subroutine mapping_apply_i
     subroutine mapping_apply_backward_i(this, dF_dX_in, dF_dX_out, X_in)
       import mapping_t, field_t
       class(mapping_t), intent(inout) :: this
       type(field_t), intent(inout) :: dF_dX_in
       type(field_t), intent(in) :: dF_dX_out, X_in
     end subroutine mapping_apply_backward_i
  end interface

  type, extends(mapping_t) :: pde_filter_t
     real(rp) :: r = real(0.01, rp)        ! the reference's default-real literal
     real(rp) :: abstol = 1e-10_rp
     integer :: max_iter = 200
     class(ax_t), allocatable :: Ax
     class(ksp_t), allocatable :: ksp
     class(pc_t), allocatable :: pc
     type(bc_list_t) :: bclst               ! all-Neumann: empty
     type(field_t) :: rhs
     type(ksp_monitor_t) :: mon
   contains
     procedure, pass(this) :: init => filter_init
     procedure, pass(this) :: apply_forward => filter_forward
     procedure, pass(this) :: apply_backward => filter_backward
     procedure, pass(this) :: free => filter_free
  end type pde_filter_t

  !> RAMP (sources/mapping_functions/RAMP_mapping.f90:72-96); defaults of
  !! :107-110.
  type, extends(mapping_t) :: ramp_mapping_t
     real(rp) :: f_min = 0.0_rp, f_max = 1000.0_rp, q = 1.0_rp
     logical :: convex_up = .true.
   contains
     procedure, pass(this) :: init => ramp_init
     procedure, pass(this) :: free => ramp_free
     procedure, pass(this) :: apply_forward => ramp_forward
     procedure, pass(this) :: apply_backward => ramp_backward
  end type ramp_mapping_t

  type :: topopt_design_t
     type(field_t) :: design_indicator, filtered_design, brinkman_amplitude, sensitivity
     type(pde_filter_t) :: filter
     type(ramp_mapping_t) :: mapping
     type(coef_t), pointer :: coef => null()
   contains
     procedure, pass(this) :: init => design_init
     procedure, pass(this) :: map_forward => design_map_forward
     procedure, pass(this) :: map_backward => design_map_backward
     procedure, pass(this) :: free => design_free
  end type topopt_design_t

contains

  subroutine mapping_init_base(this, coef)
    class(mapping_t), intent(inout) :: this
    type(coef_t), intent(inout), target :: coef
    this%coef => coef
  end subroutine mapping_init_base

  subroutine mapping_free_base(this)
    class(mapping_t), intent(inout) :: this
    this%coef => null()
  end subroutine mapping_free_base

  subroutine ramp_init(this, coef)
    class(ramp_mapping_t), intent(inout) :: this
    type(coef_t), intent(inout), target :: coef
    call this%init_base(coef)
  end subroutine ramp_init

  subroutine ramp_free(this)
    class(ramp_mapping_t), intent(inout) :: this
    call this%free_base()
  end subroutine ramp_free

  subroutine filter_init(this, coef)
    class(pde_filter_t), intent(inout) :: this
    type(coef_t), intent(inout), target :: coef
    call this%init_base(coef)
    call this%bclst%init()
    call this%rhs%init(coef%n, 'filter_rhs')
    call ax_helm_factory(this%Ax, full_formulation=.false.)
    call krylov_solver_factory(this%ksp, coef%n, 'cg', this%max_iter, this%abstol, coef)
    call precon_factory(this%pc, 'jacobi')
    call set_filter_coefficients(this)
    select type (p => this%pc)
    type is (jacobi_hip_t)
       call p%init(coef, coef%n, coef%gs_h)
    end select
    call this%ksp%set_pc(this%pc)
  end subroutine filter_init

  subroutine set_filter_coefficients(this)
    type(pde_filter_t), intent(inout) :: this
    call this%coef%h1%fill(this%r**2)
    call this%coef%h2%fill(1.0_rp)
    this%coef%ifh2 = .true.
  end subroutine set_filter_coefficients

  !> X_out = filter(X_in): RHS = gs(B X_in), solve (PDE_filter_mapping.f90:212-278)
  subroutine filter_forward(this, X_out, X_in)
    class(pde_filter_t), intent(inout) :: this
    type(field_t), intent(inout) :: X_out
    type(field_t), intent(in) :: X_in
    call set_filter_coefficients(this)
    call this%pc%update()
    call nk_check(nekoamd_col3(this%rhs%x_d, X_in%x_d, this%coef%B%x_d, this%coef%n), 'col3')
    call this%coef%gs_h%op(this%rhs, GS_OP_ADD)
    call this%bclst%apply_scalar(this%rhs%x_d, this%coef%n)
    this%mon = this%ksp%solve(this%Ax, X_out, this%rhs, this%coef%n, this%coef, &
         this%bclst, this%coef%gs_h)
  end subroutine filter_forward

  !> dF/dX_in = filter^T(dF/dX_out): RHS = B dF/dX_out, bc, gs, solve
  !! (PDE_filter_mapping.f90:298-363)
  subroutine filter_backward(this, dF_dX_in, dF_dX_out, X_in)
    class(pde_filter_t), intent(inout) :: this
    type(field_t), intent(inout) :: dF_dX_in
    type(field_t), intent(in) :: dF_dX_out, X_in
    call set_filter_coefficients(this)
    call this%pc%update()
    call nk_check(nekoamd_col3(this%rhs%x_d, dF_dX_out%x_d, this%coef%B%x_d, &
         this%coef%n), 'col3')
    call this%bclst%apply_scalar(this%rhs%x_d, this%coef%n)
    call this%coef%gs_h%op(this%rhs, GS_OP_ADD)
    this%mon = this%ksp%solve(this%Ax, dF_dX_in, this%rhs, this%coef%n, this%coef, &
         this%bclst, this%coef%gs_h)
  end subroutine filter_backward

  subroutine filter_free(this)
    class(pde_filter_t), intent(inout) :: this
    select type (k => this%ksp)
    type is (cg_hip_t)
       call k%free()
    end select
    select type (p => this%pc)
    type is (jacobi_hip_t)
       call p%free()
    end select
    call this%rhs%free()
    call this%bclst%free()
    call this%free_base()
  end subroutine filter_free

  subroutine ramp_forward(this, X_out, X_in)
    class(ramp_mapping_t), intent(inout) :: this
    type(field_t), intent(inout) :: X_out
    type(field_t), intent(in) :: X_in
    real(c_double) :: lo, hi, q
    integer(c_int) :: up
    lo = this%f_min
    hi = this%f_max
    q = this%q
    up = merge(1, 0, this%convex_up)
    call nk_check(nekoamd_ramp_fwd(X_out%x_d, X_in%x_d, lo, hi, q, up, X_in%n), 'ramp_fwd')
  end subroutine ramp_forward

  subroutine ramp_backward(this, dF_dX_in, dF_dX_out, X_in)
    class(ramp_mapping_t), intent(inout) :: this
    type(field_t), intent(inout) :: dF_dX_in
    type(field_t), intent(in) :: dF_dX_out, X_in
    real(c_double) :: lo, hi, q
    integer(c_int) :: up
    lo = this%f_min
    hi = this%f_max
    q = this%q
    up = merge(1, 0, this%convex_up)
    call nk_check(nekoamd_ramp_bwd(dF_dX_in%x_d, dF_dX_out%x_d, X_in%x_d, lo, hi, q, up, &
         X_in%n), 'ramp_bwd')
  end subroutine ramp_backward

  subroutine design_init(this, coef)
    class(topopt_design_t), intent(inout) :: this
    type(coef_t), intent(inout), target :: coef
    this%coef => coef
    call this%design_indicator%init(coef%n, 'design_indicator')
    call this%filtered_design%init(coef%n, 'filtered_design')
    call this%brinkman_amplitude%init(coef%n, 'brinkman_amplitude')
    call this%sensitivity%init(coef%n, 'sensitivity')
    call this%filter%init(coef)
    call this%mapping%init(coef)
  end subroutine design_init

  !> filter then map (topopt_design.f90:292-314; no mask in the workloads)
  subroutine design_map_forward(this)
    class(topopt_design_t), intent(inout) :: this
    call this%filter%apply_forward(this%filtered_design, this%design_indicator)
    call this%mapping%apply_forward(this%brinkman_amplitude, this%filtered_design)
  end subroutine design_map_forward

  !> chain rule (topopt_design.f90:316-352): map backward, filter backward
  subroutine design_map_backward(this, df_dchi)
    class(topopt_design_t), intent(inout) :: this
    type(field_t), intent(in) :: df_dchi
    type(field_t) :: tmp
    call tmp%init(this%coef%n, 'dF_dfiltered')
    call this%mapping%apply_backward(tmp, df_dchi, this%filtered_design)
    call this%filter%apply_backward(this%sensitivity, tmp, this%filtered_design)
    call tmp%free()
  end subroutine design_map_backward

  subroutine design_free(this)
    class(topopt_design_t), intent(inout) :: this
    call this%design_indicator%free()
    call this%filtered_design%free()
    call this%brinkman_amplitude%free()
    call this%sensitivity%free()
    call this%filter%free()
    call this%mapping%free()
  end subroutine design_free

end module nekoamd_design
