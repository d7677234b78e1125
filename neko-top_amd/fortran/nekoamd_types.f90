!> Device-resident SEM data types of the Fortran host: fields, GLL space,
!! the synthetic box mesh, geometric factors (coef_t), gather-scatter (gs_t)
!! and Dirichlet boundary lists (bc_list_t).
!!
!! They carry what Neko's field_t / space_t / dofmap_t / coef_t / gs_t /
!! bc_list_t carry on the path the reference drives
!! (sources/adjoint/adjoint_scheme.f90:331-343 builds Xh, dm_Xh, gs_Xh, c_Xh;
!! sources/adjoint/adjoint_pnpn.f90:725-726 applies gs and the bc list), but
!! every array lives in HBM: x_d, G11_d, ... are device buffers and host
!! copies are made only on request (download).  Like Neko's coef_t, coef_t
!! also carries the Helmholtz coefficients h1, h2 and the ifh2 switch that
!! ax_helm reads (sources/mapping_functions/PDE_filter_mapping.f90:225-241
!! sets them before a filter solve).
module nekoamd_types
  use, intrinsic :: iso_c_binding
  use nekoamd_capi
  implicit none
  private

  public :: rp, field_t, space_t, box_mesh_t, coef_t, gs_t, bc_list_t, &
       nk_device_init, nk_alloc, nk_free, nk_upload, nk_download, &
       GS_OP_ADD, GS_OP_MUL, GS_OP_MIN, GS_OP_MAX, NK_FACE_XMIN, NK_FACE_XMAX, &
       NK_FACE_YMIN, NK_FACE_YMAX, NK_FACE_ZMIN, NK_FACE_ZMAX

  !> A scalar field of n = lx^3 * nelv values in HBM (optional host mirror).
  type :: field_t
     integer(c_int64_t) :: n = 0
     type(c_ptr) :: x_d = c_null_ptr
     real(rp), allocatable :: x(:)
     character(len=32) :: name = ''
   contains
     procedure, pass(this) :: init => field_init
     procedure, pass(this) :: free => field_free
     procedure, pass(this) :: upload => field_upload
     procedure, pass(this) :: download => field_download
     procedure, pass(this) :: fill => field_fill
  end type field_t

  !> GLL space: nodes, weights and the derivative matrix
  !! dx(i*lx + j) = l_j'(z_i) (Neko's Xh%dxt in memory order).
  type :: space_t
     integer(c_int) :: lx = 0
     real(rp), allocatable :: zg(:), wx(:), dx(:)
     type(c_ptr) :: dx_d = c_null_ptr, zg_d = c_null_ptr, wx_d = c_null_ptr
   contains
     procedure, pass(this) :: init => space_init
     procedure, pass(this) :: free => space_free
  end type space_t

  !> nx*ny*nz hex box [0,nx/ny]x[0,1]x[0,nz/ny] with seeded vertex jitter.
  type :: box_mesh_t
     integer(c_int) :: nx = 0, ny = 0, nz = 0, nelv = 0
     real(rp) :: jitter = 0.0_rp
     integer(c_int64_t) :: seed = 0
     real(rp), allocatable :: corners(:)
   contains
     procedure, pass(this) :: init => box_init
  end type box_mesh_t

  !> Gather-scatter handle (the canonical bucketed map lives on the device).
  type :: gs_t
     type(c_ptr) :: h = c_null_ptr
   contains
     procedure, pass(this) :: init => gs_init
     procedure, pass(this) :: free => gs_free
     procedure, pass(this) :: op_field => gs_op_field
     procedure, pass(this) :: op_ptr => gs_op_ptr
     generic :: op => op_field, op_ptr
  end type gs_t

  !> Zero-Dirichlet boundary list: one 0-based int32 device mask per bc.
  type :: bc_mask_t
     type(c_ptr) :: msk_d = c_null_ptr
     integer(c_int64_t) :: nmsk = 0
  end type bc_mask_t

  type :: bc_list_t
     type(bc_mask_t), allocatable :: items(:)
     integer :: size = 0
     !> union of all masks, the form the fused CG kernels consume
     type(c_ptr) :: all_d = c_null_ptr
     integer(c_int64_t) :: nall = 0
     integer(c_int) :: face_bits = 0
   contains
     procedure, pass(this) :: init => bc_list_init
     procedure, pass(this) :: add_faces => bc_list_add_faces
     procedure, pass(this) :: apply_scalar => bc_list_apply_scalar
     procedure, pass(this) :: free => bc_list_free
  end type bc_list_t

  !> Geometric factors, mass, multiplicity and Helmholtz coefficients.
  type :: coef_t
     integer(c_int) :: lx = 0, nelv = 0
     integer(c_int64_t) :: n = 0
     type(field_t) :: G11, G12, G13, G22, G23, G33, B, jac, mult, h1, h2
     type(field_t) :: x, y, z
     logical :: ifh2 = .false.
     real(rp) :: volume = 0.0_rp
     type(space_t), pointer :: Xh => null()
     type(box_mesh_t), pointer :: msh => null()
     type(gs_t), pointer :: gs_h => null()
   contains
     procedure, pass(this) :: init => coef_init
     procedure, pass(this) :: free => coef_free
  end type coef_t

contains

  ! ---------------------------------------------------------------- device --
  subroutine nk_device_init(dev)
    integer, intent(in) :: dev
    integer(c_int) :: d
    d = dev
    call nk_check(nekoamd_device_init(d), 'device_init')
  end subroutine nk_device_init

  subroutine nk_alloc(p, bytes)
    type(c_ptr), intent(out) :: p
    integer(c_int64_t), intent(in) :: bytes
    integer(c_int64_t) :: b
    b = bytes
    call nk_check(nekoamd_device_alloc(p, b), 'device_alloc')
  end subroutine nk_alloc

  subroutine nk_free(p)
    type(c_ptr), intent(inout) :: p
    if (c_associated(p)) call nk_check(nekoamd_device_free(p), 'device_free')
    p = c_null_ptr
  end subroutine nk_free

  subroutine nk_upload(p, host, bytes)
    type(c_ptr), intent(in) :: p
    type(c_ptr), intent(in) :: host
    integer(c_int64_t), intent(in) :: bytes
    integer(c_int64_t) :: b
    b = bytes
    call nk_check(nekoamd_device_memcpy(p, host, b, NK_HOST_TO_DEVICE, 1_c_int), &
         'memcpy h2d')
  end subroutine nk_upload

  subroutine nk_download(host, p, bytes)
    type(c_ptr), intent(in) :: host
    type(c_ptr), intent(in) :: p
    integer(c_int64_t), intent(in) :: bytes
    integer(c_int64_t) :: b
    b = bytes
    call nk_check(nekoamd_device_memcpy(host, p, b, NK_DEVICE_TO_HOST, 1_c_int), &
         'memcpy d2h')
  end subroutine nk_download

  ! ----------------------------------------------------------------- field --
  subroutine field_init(this, n, name)
    class(field_t), intent(inout) :: this
    integer(c_int64_t), intent(in) :: n
    character(len=*), intent(in), optional :: name
    call this%free()
    this%n = n
    call nk_alloc(this%x_d, 8_c_int64_t * n)
    call nk_check(nekoamd_rzero(this%x_d, this%n), 'rzero')
    if (present(name)) this%name = name
  end subroutine field_init

  subroutine field_free(this)
    class(field_t), intent(inout) :: this
    call nk_free(this%x_d)
    if (allocated(this%x)) deallocate(this%x)
    this%n = 0
  end subroutine field_free

  subroutine field_upload(this, v)
    class(field_t), intent(inout) :: this
    real(rp), intent(in), target :: v(*)
    call nk_upload(this%x_d, c_loc(v), 8_c_int64_t * this%n)
  end subroutine field_upload

  subroutine field_download(this)
    class(field_t), intent(inout), target :: this
    if (.not. allocated(this%x)) allocate(this%x(this%n))
    call nk_download(c_loc(this%x), this%x_d, 8_c_int64_t * this%n)
  end subroutine field_download

  subroutine field_fill(this, c)
    class(field_t), intent(inout) :: this
    real(rp), intent(in) :: c
    real(c_double) :: cc
    cc = c
    call nk_check(nekoamd_cfill(this%x_d, cc, this%n), 'cfill')
  end subroutine field_fill

  ! ----------------------------------------------------------------- space --
  subroutine space_init(this, lx)
    class(space_t), intent(inout), target :: this
    integer, intent(in) :: lx
    integer(c_int) :: l
    call this%free()
    this%lx = lx
    l = lx
    allocate(this%zg(lx), this%wx(lx), this%dx(lx * lx))
    call nk_check(nekoamd_gll(l, this%zg, this%wx, this%dx), 'gll')
    call nk_alloc(this%dx_d, int(8 * lx * lx, c_int64_t))
    call nk_alloc(this%zg_d, int(8 * lx, c_int64_t))
    call nk_alloc(this%wx_d, int(8 * lx, c_int64_t))
    call nk_upload(this%dx_d, c_loc(this%dx), int(8 * lx * lx, c_int64_t))
    call nk_upload(this%zg_d, c_loc(this%zg), int(8 * lx, c_int64_t))
    call nk_upload(this%wx_d, c_loc(this%wx), int(8 * lx, c_int64_t))
  end subroutine space_init

  subroutine space_free(this)
    class(space_t), intent(inout) :: this
    call nk_free(this%dx_d)
    call nk_free(this%zg_d)
    call nk_free(this%wx_d)
    if (allocated(this%zg)) deallocate(this%zg, this%wx, this%dx)
  end subroutine space_free

  ! ------------------------------------------------------------------ mesh --
  subroutine box_init(this, nx, ny, nz, jitter, seed)
    class(box_mesh_t), intent(inout) :: this
    integer, intent(in) :: nx, ny, nz
    real(rp), intent(in) :: jitter
    integer(c_int64_t), intent(in) :: seed
    real(c_double), allocatable :: v(:)
    real(c_double) :: lxyz(3), jit
    integer(c_int64_t) :: s
    this%nx = nx
    this%ny = ny
    this%nz = nz
    this%nelv = nx * ny * nz
    this%jitter = jitter
    this%seed = seed
    lxyz = [real(nx, rp) / ny, 1.0_rp, real(nz, rp) / ny]
    jit = jitter
    s = seed
    allocate(v(3 * (nx + 1) * (ny + 1) * (nz + 1)))
    if (allocated(this%corners)) deallocate(this%corners)
    allocate(this%corners(24 * this%nelv))
    call nk_check(nekoamd_box_vertices(this%nx, this%ny, this%nz, lxyz, jit, s, v), &
         'box_vertices')
    call nk_check(nekoamd_box_corners(this%nx, this%ny, this%nz, v, this%corners), &
         'box_corners')
  end subroutine box_init

  ! -------------------------------------------------------------------- gs --
  subroutine gs_init(this, gid, n)
    class(gs_t), intent(inout) :: this
    integer(c_int64_t), intent(in) :: n
    integer(c_int64_t), intent(in) :: gid(n)
    integer(c_int64_t) :: nn
    integer(c_int64_t), allocatable :: g(:)
    nn = n
    g = gid
    call nk_check(nekoamd_gs_init(this%h, g, nn), 'gs_init')
  end subroutine gs_init

  subroutine gs_free(this)
    class(gs_t), intent(inout) :: this
    if (c_associated(this%h)) call nk_check(nekoamd_gs_free(this%h), 'gs_free')
    this%h = c_null_ptr
  end subroutine gs_free

  !> gs_t%op(field, op)  (sources/adjoint/adjoint_pnpn.f90:725)
  subroutine gs_op_field(this, f, op)
    class(gs_t), intent(inout) :: this
    type(field_t), intent(inout) :: f
    integer(c_int), intent(in) :: op
    integer(c_int) :: o
    o = op
    call nk_check(nekoamd_gs_op_f64(this%h, f%x_d, o), 'gs_op')
  end subroutine gs_op_field

  !> gs_t%op(u_d, n, op) on a raw device buffer
  subroutine gs_op_ptr(this, u_d, n, op)
    class(gs_t), intent(inout) :: this
    type(c_ptr), intent(in) :: u_d
    integer(c_int64_t), intent(in) :: n
    integer(c_int), intent(in) :: op
    integer(c_int) :: o
    o = op
    if (n .lt. 0) error stop 'gs_op: negative size'
    call nk_check(nekoamd_gs_op_f64(this%h, u_d, o), 'gs_op')
  end subroutine gs_op_ptr

  ! ------------------------------------------------------------------- bcs --
  subroutine bc_list_init(this)
    class(bc_list_t), intent(inout) :: this
    call this%free()
    allocate(this%items(0))
    this%size = 0
  end subroutine bc_list_init

  !> Add the zero-Dirichlet points on the given box faces as one bc.
  subroutine bc_list_add_faces(this, msh, lx, faces)
    class(bc_list_t), intent(inout) :: this
    type(box_mesh_t), intent(in) :: msh
    integer, intent(in) :: lx, faces
    type(bc_mask_t), allocatable :: tmp(:)
    integer(c_int) :: l, f
    integer(c_int64_t) :: cnt, all_cnt
    integer(c_int32_t), allocatable, target :: m(:), all(:)
    l = lx
    f = faces
    call nk_check(nekoamd_box_face_mask(msh%nx, msh%ny, msh%nz, l, f, c_null_ptr, cnt), &
         'face_mask')
    allocate(m(max(cnt, 1_c_int64_t)))
    call nk_check(nekoamd_box_face_mask(msh%nx, msh%ny, msh%nz, l, f, c_loc(m), cnt), &
         'face_mask')
    allocate(tmp(this%size + 1))
    if (this%size .gt. 0) tmp(1:this%size) = this%items
    tmp(this%size + 1)%nmsk = cnt
    if (cnt .gt. 0) then
       call nk_alloc(tmp(this%size + 1)%msk_d, 4_c_int64_t * cnt)
       call nk_upload(tmp(this%size + 1)%msk_d, c_loc(m), 4_c_int64_t * cnt)
    end if
    call move_alloc(tmp, this%items)
    this%size = this%size + 1
    ! union of all face masks: recompute from the OR of all faces added
    call nk_free(this%all_d)
    this%nall = 0
    f = ior(int(faces, c_int), this%face_bits)
    call nk_check(nekoamd_box_face_mask(msh%nx, msh%ny, msh%nz, l, f, c_null_ptr, &
         all_cnt), 'face_mask')
    allocate(all(max(all_cnt, 1_c_int64_t)))
    call nk_check(nekoamd_box_face_mask(msh%nx, msh%ny, msh%nz, l, f, c_loc(all), &
         all_cnt), 'face_mask')
    this%nall = all_cnt
    if (all_cnt .gt. 0) then
       call nk_alloc(this%all_d, 4_c_int64_t * all_cnt)
       call nk_upload(this%all_d, c_loc(all), 4_c_int64_t * all_cnt)
    end if
    this%face_bits = f
  end subroutine bc_list_add_faces

  !> bc_list_apply_scalar(bclst, x, n): zero x on every Dirichlet point
  !! (sources/adjoint/adjoint_pnpn.f90:726).
  subroutine bc_list_apply_scalar(this, x_d, n)
    class(bc_list_t), intent(inout) :: this
    type(c_ptr), intent(in) :: x_d
    integer(c_int64_t), intent(in) :: n
    integer :: i
    if (n .lt. 0) error stop 'bc_list_apply_scalar: negative size'
    do i = 1, this%size
       if (this%items(i)%nmsk .gt. 0) &
            call nk_check(nekoamd_bc_zero(x_d, this%items(i)%msk_d, this%items(i)%nmsk), &
            'bc_zero')
    end do
  end subroutine bc_list_apply_scalar

  subroutine bc_list_free(this)
    class(bc_list_t), intent(inout) :: this
    integer :: i
    if (allocated(this%items)) then
       do i = 1, this%size
          call nk_free(this%items(i)%msk_d)
       end do
       deallocate(this%items)
    end if
    call nk_free(this%all_d)
    this%size = 0
    this%nall = 0
    this%face_bits = 0
  end subroutine bc_list_free

  ! ------------------------------------------------------------------ coef --
  !> Geometric factors on the device, mult = 1/gs(1), volume = sum B.
  subroutine coef_init(this, Xh, msh, gs_h)
    class(coef_t), intent(inout) :: this
    type(space_t), intent(in), target :: Xh
    type(box_mesh_t), intent(in), target :: msh
    type(gs_t), intent(inout), target :: gs_h
    type(c_ptr) :: corners_d
    type(field_t) :: ones
    integer(c_int64_t) :: n
    integer(c_int) :: lx, nelv
    real(rp), allocatable, target :: cor(:)
    real(c_double) :: vol
    lx = Xh%lx
    nelv = msh%nelv
    n = int(nelv, c_int64_t) * lx * lx * lx
    this%lx = lx
    this%nelv = nelv
    this%n = n
    this%Xh => Xh
    this%msh => msh
    this%gs_h => gs_h
    call this%G11%init(n, 'G11')
    call this%G12%init(n, 'G12')
    call this%G13%init(n, 'G13')
    call this%G22%init(n, 'G22')
    call this%G23%init(n, 'G23')
    call this%G33%init(n, 'G33')
    call this%B%init(n, 'B')
    call this%jac%init(n, 'jac')
    call this%x%init(n, 'x')
    call this%y%init(n, 'y')
    call this%z%init(n, 'z')
    call this%mult%init(n, 'mult')
    call this%h1%init(n, 'h1')
    call this%h2%init(n, 'h2')
    call this%h1%fill(1.0_rp)
    cor = msh%corners
    call nk_alloc(corners_d, int(8 * size(cor), c_int64_t))
    call nk_upload(corners_d, c_loc(cor), int(8 * size(cor), c_int64_t))
    call nk_check(nekoamd_coef_geo(this%x%x_d, this%y%x_d, this%z%x_d, this%G11%x_d, &
         this%G12%x_d, this%G13%x_d, this%G22%x_d, this%G23%x_d, this%G33%x_d, &
         this%B%x_d, this%jac%x_d, corners_d, Xh%zg_d, Xh%wx_d, Xh%dx_d, lx, nelv), &
         'coef_geo')
    call nk_free(corners_d)
    call this%mult%fill(1.0_rp)
    call gs_h%op(this%mult, GS_OP_ADD)
    call nk_check(nekoamd_invcol1(this%mult%x_d, n), 'invcol1')
    call ones%init(n)
    call ones%fill(1.0_rp)
    call nk_check(nekoamd_glsc2(vol, this%B%x_d, ones%x_d, n), 'glsc2')
    this%volume = vol
    call ones%free()
  end subroutine coef_init

  subroutine coef_free(this)
    class(coef_t), intent(inout) :: this
    call this%G11%free()
    call this%G12%free()
    call this%G13%free()
    call this%G22%free()
    call this%G23%free()
    call this%G33%free()
    call this%B%free()
    call this%jac%free()
    call this%x%free()
    call this%y%free()
    call this%z%free()
    call this%mult%free()
    call this%h1%free()
    call this%h2%free()
  end subroutine coef_free

end module nekoamd_types
