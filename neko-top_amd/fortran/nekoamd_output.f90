!> Neko's fld_file_output_t as the reference uses it
!! (steady_state_problem.f90:228-244, topopt_design.f90:283-286,
!! rugby_ball/driver.f90:143): a named series of nfields device fields;
!! every sample(t) writes <name>0.f<counter, 5 digits> through
!! nekoamd_fld_write (the mesh into the first file only) and rewrites the
!! series index <name>.nek5000.
module nekoamd_output
  use, intrinsic :: iso_c_binding
  use nekoamd_capi
  use nekoamd_types
  implicit none
  private

  public :: fld_file_output_t

  type :: fld_file_output_t
     character(len=:), allocatable :: name
     integer :: counter = 0
     integer(c_int) :: wdsize = 4
     type(c_ptr), allocatable :: fields(:)
     type(coef_t), pointer :: coef => null()
   contains
     procedure, pass(this) :: init => fld_output_init
     procedure, pass(this) :: assign => fld_output_assign
     procedure, pass(this) :: sample => fld_output_sample
     procedure, pass(this) :: free => fld_output_free
  end type fld_file_output_t

contains

  !> wdsize: 4 (Neko's default single precision) or 8
  subroutine fld_output_init(this, coef, name, nfields, wdsize)
    class(fld_file_output_t), intent(inout) :: this
    type(coef_t), intent(in), target :: coef
    character(len=*), intent(in) :: name
    integer, intent(in) :: nfields, wdsize
    call this%free()
    this%coef => coef
    this%name = trim(name)
    this%wdsize = int(wdsize, c_int)
    allocate(this%fields(nfields))
    this%fields = c_null_ptr
    this%counter = 0
  end subroutine fld_output_init

  subroutine fld_output_assign(this, i, f)
    class(fld_file_output_t), intent(inout) :: this
    integer, intent(in) :: i
    type(field_t), intent(in) :: f
    this%fields(i) = f%x_d
  end subroutine fld_output_assign

  subroutine fld_output_sample(this, t)
    class(fld_file_output_t), intent(inout) :: this
    real(rp), intent(in) :: t
    character(len=:), allocatable, target :: path, index, prefix
    character(len=5) :: num
    real(c_double) :: tt
    integer(c_int) :: step, nf, dev, first, count
    type(c_ptr), target, allocatable :: ptrs(:)
    type(c_ptr) :: x, y, z
    write(num, '(I5.5)') this%counter
    path = this%name // '0.f' // num // c_null_char
    index = this%name // '.nek5000' // c_null_char
    prefix = this%name // c_null_char
    tt = t
    step = int(this%counter, c_int)
    nf = int(size(this%fields), c_int)
    dev = 1
    allocate(ptrs(max(nf, 1)))
    if (nf > 0) ptrs(1:nf) = this%fields
    x = c_null_ptr
    y = c_null_ptr
    z = c_null_ptr
    if (this%counter .eq. 0) then
       x = this%coef%x%x_d
       y = this%coef%y%x_d
       z = this%coef%z%x_d
    end if
    call nk_check(nekoamd_fld_write(c_loc(path), tt, step, this%coef%lx, this%coef%nelv, &
         nf, c_loc(ptrs), x, y, z, this%wdsize, dev), 'fld_write')
    first = 0
    count = int(this%counter + 1, c_int)
    call nk_check(nekoamd_fld_series(c_loc(index), c_loc(prefix), first, count), 'fld_series')
    this%counter = this%counter + 1
  end subroutine fld_output_sample

  subroutine fld_output_free(this)
    class(fld_file_output_t), intent(inout) :: this
    if (allocated(this%fields)) deallocate(this%fields)
    this%coef => null()
    this%counter = 0
  end subroutine fld_output_free

end module nekoamd_output
