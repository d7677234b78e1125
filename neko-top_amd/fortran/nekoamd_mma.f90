!> mma_t: the reference's method-of-moving-asymptotes optimiser
!! (sources/mma/mma.f90:49-98) over the device MMA of libnekoamd: init
!! (mma_init_attributes, :186-295, with the defaults of mma_init_json,
!! :160-176), update (mma_update_vector, :326-353: gensub, then the
!! dual-primal subsolve), KKT (:355-393), get_n / get_m / get_residumax /
!! get_residunorm (:434-456) and free (:396-429).
!!
!! The design variables, bounds and gradients are device fields (field_t);
!! dfdx is m x n, column-major like the reference's matrix_t (for m = 1 a
!! plain field).  fval, a, c, d are host arrays of m.  Like the reference,
!! residumax and residunorm start at 1e5 and are set by KKT.
module nekoamd_mma
  use, intrinsic :: iso_c_binding
  use nekoamd_capi
  use nekoamd_types
  implicit none
  private

  public :: mma_t

  type :: mma_t
     private
     type(c_ptr) :: h = c_null_ptr
     integer :: n = 0, m = 0
     real(rp) :: residumax = 1.0e5_rp, residunorm = 1.0e5_rp
     integer :: newton_iters = 0
   contains
     procedure, public, pass(this) :: init => mma_init_attributes
     procedure, public, pass(this) :: free => mma_free
     procedure, public, pass(this) :: update => mma_update
     procedure, public, pass(this) :: KKT => mma_kkt
     procedure, public, pass(this) :: get_n => mma_get_n
     procedure, public, pass(this) :: get_m => mma_get_m
     procedure, public, pass(this) :: get_residumax => mma_get_residumax
     procedure, public, pass(this) :: get_residunorm => mma_get_residunorm
     procedure, public, pass(this) :: get_newton_iters => mma_get_newton_iters
  end type mma_t

contains

  !> x: the initial design (its values seed xold1/xold2); n design
  !! variables, m constraints; a0, a, c, d the MMA problem constants; xmin,
  !! xmax the bounds.  Optional: max_iter (100), epsimin
  !! (1e-9 sqrt(m + n)), asyinit (0.5), asyincr (1.2), asydecr (0.7).
  subroutine mma_init_attributes(this, x, n, m, a0, a, c, d, xmin, xmax, max_iter, &
       epsimin, asyinit, asyincr, asydecr)
    class(mma_t), intent(inout) :: this
    type(field_t), intent(in) :: x, xmin, xmax
    integer, intent(in) :: n, m
    real(rp), intent(in) :: a0, a(m), c(m), d(m)
    integer, intent(in), optional :: max_iter
    real(rp), intent(in), optional :: epsimin, asyinit, asyincr, asydecr
    integer(c_int) :: nn, mm, mi
    real(c_double) :: a0c, ac(m), cc(m), dc(m), eps, ai, au, ad
    call this%free()
    if (n .le. 0 .or. m .lt. 1) error stop 'mma_t%init: need n > 0 and m >= 1'
    if (x%n .ne. n .or. xmin%n .ne. n .or. xmax%n .ne. n) &
         error stop 'mma_t%init: x, xmin, xmax must hold n values'
    nn = n
    mm = m
    a0c = a0
    ac = a
    cc = c
    dc = d
    mi = 100
    if (present(max_iter)) mi = max_iter
    eps = 1.0e-9_rp * sqrt(real(m + n, rp))
    if (present(epsimin)) eps = epsimin
    ai = 0.5_rp
    if (present(asyinit)) ai = asyinit
    au = 1.2_rp
    if (present(asyincr)) au = asyincr
    ad = 0.7_rp
    if (present(asydecr)) ad = asydecr
    call nk_check(nekoamd_mma_create(this%h, nn, mm, x%x_d, a0c, ac, cc, dc, xmin%x_d, &
         xmax%x_d, mi, eps, ai, au, ad), 'mma_create')
    this%n = n
    this%m = m
    this%residumax = 1.0e5_rp
    this%residunorm = 1.0e5_rp
  end subroutine mma_init_attributes

  subroutine mma_free(this)
    class(mma_t), intent(inout) :: this
    if (c_associated(this%h)) call nk_check(nekoamd_mma_free(this%h), 'mma_free')
    this%h = c_null_ptr
    this%n = 0
    this%m = 0
  end subroutine mma_free

  !> One MMA iteration: x is overwritten with the new design.
  subroutine mma_update(this, iter, x, df0dx, fval, dfdx)
    class(mma_t), intent(inout) :: this
    integer, intent(in) :: iter
    type(field_t), intent(inout) :: x
    type(field_t), intent(in) :: df0dx, dfdx
    real(rp), intent(in) :: fval(this%m)
    integer(c_int) :: it
    real(c_double) :: fv(this%m)
    if (.not. c_associated(this%h)) error stop 'mma_t%update: not initialised'
    it = iter
    fv = fval
    call nk_check(nekoamd_mma_update(this%h, it, x%x_d, df0dx%x_d, fv, dfdx%x_d), &
         'mma_update')
  end subroutine mma_update

  !> KKT residual of the original problem at x (with the duals of the last
  !! update): sets residumax and residunorm.
  subroutine mma_kkt(this, x, df0dx, fval, dfdx)
    class(mma_t), intent(inout) :: this
    type(field_t), intent(in) :: x, df0dx, dfdx
    real(rp), intent(in) :: fval(this%m)
    real(c_double) :: fv(this%m), res(3)
    if (.not. c_associated(this%h)) error stop 'mma_t%KKT: not initialised'
    fv = fval
    call nk_check(nekoamd_mma_kkt(this%h, x%x_d, df0dx%x_d, fv, dfdx%x_d), 'mma_kkt')
    call nk_check(nekoamd_mma_residuals(this%h, res), 'mma_residuals')
    this%residumax = res(1)
    this%residunorm = res(2)
    this%newton_iters = int(res(3))
  end subroutine mma_kkt

  pure function mma_get_n(this) result(n)
    class(mma_t), intent(in) :: this
    integer :: n
    n = this%n
  end function mma_get_n

  pure function mma_get_m(this) result(m)
    class(mma_t), intent(in) :: this
    integer :: m
    m = this%m
  end function mma_get_m

  pure function mma_get_residumax(this) result(r)
    class(mma_t), intent(in) :: this
    real(rp) :: r
    r = this%residumax
  end function mma_get_residumax

  pure function mma_get_residunorm(this) result(r)
    class(mma_t), intent(in) :: this
    real(rp) :: r
    r = this%residunorm
  end function mma_get_residunorm

  !> Newton iterations of the last update's sub-solve (no reference
  !! counterpart; reported by the optimisation loop).
  pure function mma_get_newton_iters(this) result(k)
    class(mma_t), intent(in) :: this
    integer :: k
    k = this%newton_iters
  end function mma_get_newton_iters

end module nekoamd_mma
