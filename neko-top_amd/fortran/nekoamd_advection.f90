!> The advection terms of the Pn/Pn step, with or without dealiasing:
!! neko-top's advection_adjoint_t family (sources/adjoint/advection_adjoint.f90)
!! with its two members adv_lin_no_dealias_t (adv_adjoint_no_dealias.f90) and
!! adv_lin_dealias_t (adv_adjoint_dealias.f90), built by
!! advection_adjoint_factory (advection_adjoint_fctry.f90:57-99), plus the
!! forward convective term of Neko's advection_t%compute.
!!   compute          f = -B (u . grad) u (collocated), or
!!                    f = -J^T [ t . opgrad(t_c) ] on lxd GL points
!!   compute_adjoint  f -= adjoint advection of u about the base flow ub
!!   compute_linear   f -= u' . grad U + U . grad u'  (dealiased form only)
!! compute overwrites f; compute_adjoint and compute_linear subtract from it,
!! as the reference's do.
module nekoamd_advection
  use, intrinsic :: iso_c_binding
  use nekoamd_capi
  use nekoamd_types
  implicit none
  private

  public :: advection_adjoint_t, advection_adjoint_factory

  type :: advection_adjoint_t
     logical :: dealias = .false.
     integer(c_int) :: lxd = 0
     !> dealiased: the GL rule, interpolator and GL metric (coef_GL)
     type(c_ptr) :: state = c_null_ptr
     !> collocated adjoint: 12 n doubles of work, allocated on first use
     type(c_ptr) :: work = c_null_ptr
     type(coef_t), pointer :: coef => null()
   contains
     procedure, pass(this) :: compute => adv_compute
     procedure, pass(this) :: compute_adjoint => adv_compute_adjoint
     procedure, pass(this) :: compute_linear => adv_compute_linear
     procedure, pass(this) :: free => adv_free
  end type advection_adjoint_t

contains

  !> dealias and lxd as case.numerics.dealias and
  !! case.numerics.dealiased_polynomial_order; lxd <= 0 takes the
  !! reference's default 3 (p + 1) / 2 (advection_adjoint_fctry.f90:66-70).
  subroutine advection_adjoint_factory(object, coef, dealias, lxd)
    type(advection_adjoint_t), intent(inout) :: object
    type(coef_t), intent(inout), target :: coef
    logical, intent(in) :: dealias
    integer, intent(in) :: lxd
    call object%free()
    object%coef => coef
    object%dealias = dealias
    if (dealias) then
       object%lxd = lxd
       if (lxd .le. 0) object%lxd = 3 * coef%lx / 2
       call nk_check(nekoamd_dealias_create(object%state, coef%x%x_d, coef%y%x_d, &
            coef%z%x_d, coef%Xh%dx_d, coef%lx, object%lxd, coef%nelv), 'dealias_create')
    end if
  end subroutine advection_adjoint_factory

  subroutine adv_compute(this, u, v, w, fx, fy, fz)
    class(advection_adjoint_t), intent(inout) :: this
    type(field_t), intent(inout) :: u, v, w, fx, fy, fz
    real(c_double) :: beta
    associate(c => this%coef)
      if (this%dealias) then
         beta = 0.0_c_double
         call nk_check(nekoamd_dealias_advection(this%state, fx%x_d, fy%x_d, fz%x_d, &
              u%x_d, v%x_d, w%x_d, beta), 'dealias_advection')
      else
         call nk_check(nekoamd_advection(fx%x_d, fy%x_d, fz%x_d, u%x_d, v%x_d, w%x_d, &
              c%B%x_d, c%x%x_d, c%y%x_d, c%z%x_d, c%Xh%dx_d, c%lx, c%nelv), 'advection')
      end if
    end associate
  end subroutine adv_compute

  subroutine adv_compute_adjoint(this, vx, vy, vz, vxb, vyb, vzb, fx, fy, fz)
    class(advection_adjoint_t), intent(inout) :: this
    type(field_t), intent(inout) :: vx, vy, vz, vxb, vyb, vzb, fx, fy, fz
    associate(c => this%coef)
      if (this%dealias) then
         call nk_check(nekoamd_dealias_adjoint(this%state, fx%x_d, fy%x_d, fz%x_d, &
              vx%x_d, vy%x_d, vz%x_d, vxb%x_d, vyb%x_d, vzb%x_d), 'dealias_adjoint')
      else
         if (.not. c_associated(this%work)) call nk_alloc(this%work, 8_c_int64_t * 12 * c%n)
         call nk_check(nekoamd_adjoint_advection(fx%x_d, fy%x_d, fz%x_d, vx%x_d, vy%x_d, &
              vz%x_d, vxb%x_d, vyb%x_d, vzb%x_d, c%x%x_d, c%y%x_d, c%z%x_d, c%Xh%dx_d, &
              c%Xh%wx_d, c%lx, c%nelv, this%work), 'adjoint_advection')
      end if
    end associate
  end subroutine adv_compute_adjoint

  subroutine adv_compute_linear(this, vx, vy, vz, vxb, vyb, vzb, fx, fy, fz)
    class(advection_adjoint_t), intent(inout) :: this
    type(field_t), intent(inout) :: vx, vy, vz, vxb, vyb, vzb, fx, fy, fz
    if (.not. this%dealias) then
       write(*, '(A)') 'nekoamd: compute_linear is built for the dealiased form only'
       error stop 'nekoamd: unsupported advection form'
    end if
    call nk_check(nekoamd_dealias_linear(this%state, fx%x_d, fy%x_d, fz%x_d, vx%x_d, &
         vy%x_d, vz%x_d, vxb%x_d, vyb%x_d, vzb%x_d), 'dealias_linear')
  end subroutine adv_compute_linear

  subroutine adv_free(this)
    class(advection_adjoint_t), intent(inout) :: this
    if (c_associated(this%state)) call nk_check(nekoamd_dealias_free(this%state), &
         'dealias_free')
    this%state = c_null_ptr
    call nk_free(this%work)
    this%dealias = .false.
    this%coef => null()
  end subroutine adv_free

end module nekoamd_advection
