!> Source terms of the momentum equation and their handler:
!!
!!   source_term_t          the abstract term (init_base, compute_, free:
!!                          sources/source_terms/simple_brinkman_source_term.f90:52-67,
!!                          :118 extends Neko's source_term_t)
!!   source_term_handler_t  add(term) (sources/problems/steady_state_problem.f90:153,
!!                          :163) and compute(t, tstep) (sources/adjoint/adjoint_pnpn.f90:669)
!!   simple_brinkman_source_term_t
!!                          f -= chi u (simple_brinkman_source_term.f90:139-153)
!!
!! Two departures, both for the device step: a term's compute_ adds its
!! contribution times the mass matrix B (the reference multiplies the
!! handler's sum by B afterwards, opcolv at adjoint_pnpn.f90:672-676), and
!! the handler adds to the right-hand side instead of zeroing it first, so
!! the step can write the advection term before it (the sum is the same).
!!
!! A term whose contribution is a sink -sigma u can instead be treated
!! semi-implicitly: it hands sigma to the step (implicit_sink), which puts
!! it into the velocity Helmholtz operator (h2 += sigma) and sigma B u_e
!! into the pressure forcing.  The Brinkman term does this by default; its
!! explicit mode is the reference's treatment (stable only for
!! chi dt < 4/3 with BDF2/EXT2).
module nekoamd_source_term
  use, intrinsic :: iso_c_binding
  use nekoamd_capi
  use nekoamd_types
  implicit none
  private

  public :: source_term_t, source_term_handler_t, simple_brinkman_source_term_t

  type, abstract :: source_term_t
     !> the right-hand side the term adds to (Neko's fields f_x, f_y, f_z)
     type(field_t), pointer :: f_x => null(), f_y => null(), f_z => null()
     type(coef_t), pointer :: coef => null()
     !> the term acts for start_time <= t <= end_time
     real(rp) :: start_time = 0.0_rp, end_time = huge(0.0_rp)
   contains
     procedure, pass(this) :: init_base => source_term_init_base
     procedure, pass(this) :: free_base => source_term_free_base
     procedure, pass(this) :: compute => source_term_compute
     procedure, pass(this) :: implicit_sink => source_term_no_sink
     procedure(source_term_compute_i), pass(this), deferred :: compute_
     procedure(source_term_free_i), pass(this), deferred :: free
  end type source_term_t

  abstract interface
     subroutine source_term_compute_i(this, t, tstep)
       import source_term_t, rp
       class(source_term_t), intent(inout) :: this
       real(rp), intent(in) :: t
       integer, intent(in) :: tstep
     end subroutine source_term_compute_i
     subroutine source_term_free_i(this)
       import source_term_t
       class(source_term_t), intent(inout) :: this
     end subroutine source_term_free_i
  end interface

  type :: source_term_ptr_t
     class(source_term_t), pointer :: p => null()
  end type source_term_ptr_t

  type :: source_term_handler_t
     type(source_term_ptr_t), allocatable :: items(:)
     integer :: size = 0
     !> sum of the implicit sinks (more than one term), or zero (none)
     type(field_t) :: sink, zero
     integer(c_int64_t) :: n = 0
   contains
     procedure, pass(this) :: init => handler_init
     procedure, pass(this) :: add => handler_add
     procedure, pass(this) :: compute => handler_compute
     procedure, pass(this) :: implicit_sink => handler_implicit_sink
     procedure, pass(this) :: free => handler_free
  end type source_term_handler_t

  !> Brinkman penalisation f -= chi u, chi the design's Brinkman amplitude,
  !! u the velocity the solve advances (the primal in the forward solve, the
  !! adjoint velocity in the adjoint solve, steady_state_problem.f90:147-163).
  type, extends(source_term_t) :: simple_brinkman_source_term_t
     type(field_t), pointer :: chi => null()
     type(field_t), pointer :: u => null(), v => null(), w => null()
     logical :: implicit = .true.
   contains
     procedure, pass(this) :: init => brinkman_init
     procedure, pass(this) :: compute_ => brinkman_compute
     procedure, pass(this) :: implicit_sink => brinkman_sink
     procedure, pass(this) :: free => brinkman_free
  end type simple_brinkman_source_term_t

contains

  subroutine source_term_init_base(this, f_x, f_y, f_z, coef, start_time, end_time)
    class(source_term_t), intent(inout) :: this
    type(field_t), intent(inout), target :: f_x, f_y, f_z
    type(coef_t), intent(inout), target :: coef
    real(rp), intent(in), optional :: start_time, end_time
    this%f_x => f_x
    this%f_y => f_y
    this%f_z => f_z
    this%coef => coef
    if (present(start_time)) this%start_time = start_time
    if (present(end_time)) this%end_time = end_time
  end subroutine source_term_init_base

  subroutine source_term_free_base(this)
    class(source_term_t), intent(inout) :: this
    this%f_x => null()
    this%f_y => null()
    this%f_z => null()
    this%coef => null()
  end subroutine source_term_free_base

  !> compute_ inside the term's time window
  subroutine source_term_compute(this, t, tstep)
    class(source_term_t), intent(inout) :: this
    real(rp), intent(in) :: t
    integer, intent(in) :: tstep
    if (t .ge. this%start_time .and. t .le. this%end_time) call this%compute_(t, tstep)
  end subroutine source_term_compute

  !> default: no implicit part
  function source_term_no_sink(this, t) result(p)
    class(source_term_t), intent(inout) :: this
    real(rp), intent(in) :: t
    type(c_ptr) :: p
    p = c_null_ptr
  end function source_term_no_sink

  ! ------------------------------------------------------------- handler --
  subroutine handler_init(this, n)
    class(source_term_handler_t), intent(inout) :: this
    integer(c_int64_t), intent(in) :: n
    call this%free()
    allocate(this%items(0))
    this%size = 0
    this%n = n
    call this%zero%init(n, 'no_sink')
  end subroutine handler_init

  !> add(term): the handler keeps a pointer; the term must outlive it
  subroutine handler_add(this, term)
    class(source_term_handler_t), intent(inout) :: this
    class(source_term_t), intent(inout), target :: term
    type(source_term_ptr_t), allocatable :: tmp(:)
    allocate(tmp(this%size + 1))
    if (this%size .gt. 0) tmp(1:this%size) = this%items(1:this%size)
    tmp(this%size + 1)%p => term
    call move_alloc(tmp, this%items)
    this%size = this%size + 1
  end subroutine handler_add

  !> every term's explicit contribution (times B) added to its right-hand side
  subroutine handler_compute(this, t, tstep)
    class(source_term_handler_t), intent(inout) :: this
    real(rp), intent(in) :: t
    integer, intent(in) :: tstep
    integer :: i
    do i = 1, this%size
       call this%items(i)%p%compute(t, tstep)
    end do
  end subroutine handler_compute

  !> Device pointer of the total implicit sink sigma at time t: the single
  !! term's own field when one term has one, their sum when several do, a
  !! zero field when none does.
  function handler_implicit_sink(this, t) result(p)
    class(source_term_handler_t), intent(inout) :: this
    real(rp), intent(in) :: t
    type(c_ptr) :: p, q
    integer :: i, k
    p = this%zero%x_d
    k = 0
    do i = 1, this%size
       associate(term => this%items(i)%p)
         if (t .lt. term%start_time .or. t .gt. term%end_time) cycle
         q = term%implicit_sink(t)
       end associate
       if (.not. c_associated(q)) cycle
       k = k + 1
       if (k .eq. 1) then
          p = q
       else
          if (k .eq. 2) then
             if (this%sink%n .ne. this%n) call this%sink%init(this%n, 'sink')
             call nk_check(nekoamd_copy(this%sink%x_d, p, this%n), 'copy')
             p = this%sink%x_d
          end if
          call nk_check(nekoamd_add2(this%sink%x_d, q, this%n), 'add2')
       end if
    end do
  end function handler_implicit_sink

  subroutine handler_free(this)
    class(source_term_handler_t), intent(inout) :: this
    if (allocated(this%items)) deallocate(this%items)
    this%size = 0
    call this%sink%free()
    call this%zero%free()
    this%n = 0
  end subroutine handler_free

  ! ------------------------------------------------------------ Brinkman --
  !> f_x/f_y/f_z: the right-hand side; u, v, w: the velocity it acts on;
  !! chi: the Brinkman amplitude; implicit: semi-implicit (default) or the
  !! reference's explicit treatment.
  subroutine brinkman_init(this, f_x, f_y, f_z, coef, chi, u, v, w, implicit, &
       start_time, end_time)
    class(simple_brinkman_source_term_t), intent(inout) :: this
    type(field_t), intent(inout), target :: f_x, f_y, f_z, chi, u, v, w
    type(coef_t), intent(inout), target :: coef
    logical, intent(in), optional :: implicit
    real(rp), intent(in), optional :: start_time, end_time
    call this%init_base(f_x, f_y, f_z, coef, start_time, end_time)
    this%chi => chi
    this%u => u
    this%v => v
    this%w => w
    this%implicit = .true.
    if (present(implicit)) this%implicit = implicit
  end subroutine brinkman_init

  !> explicit mode: f -= B chi u (field_subcol3 and the mass in one pass)
  subroutine brinkman_compute(this, t, tstep)
    class(simple_brinkman_source_term_t), intent(inout) :: this
    real(rp), intent(in) :: t
    integer, intent(in) :: tstep
    if (this%implicit) return
    call nk_check(nekoamd_brinkman_mass(this%f_x%x_d, this%f_y%x_d, this%f_z%x_d, &
         this%u%x_d, this%v%x_d, this%w%x_d, this%chi%x_d, this%coef%B%x_d, &
         this%coef%n), 'brinkman_mass')
  end subroutine brinkman_compute

  function brinkman_sink(this, t) result(p)
    class(simple_brinkman_source_term_t), intent(inout) :: this
    real(rp), intent(in) :: t
    type(c_ptr) :: p
    p = c_null_ptr
    if (this%implicit) p = this%chi%x_d
  end function brinkman_sink

  subroutine brinkman_free(this)
    class(simple_brinkman_source_term_t), intent(inout) :: this
    this%chi => null()
    this%u => null()
    this%v => null()
    this%w => null()
    call this%free_base()
  end subroutine brinkman_free

end module nekoamd_source_term
