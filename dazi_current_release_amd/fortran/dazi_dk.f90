! dazi_dk.f90 -- Fortran binding of the MI355X depth-kernel stage (C ABI in
! include/dazi_dk.h, library libdazi_dk.so).
!
! depthkernel_hip and depthkernelTI_hip take exactly the argument lists of the
! reference subroutines
!   depthkernel   code/src_inv_iso_joint/CalSurfG.f90:1-15
!   depthkernelTI code/src_inv_iso_joint/depthkernelTI.f90:2-34
! so the reference's backend dispatch (CalSurfG.f90:37-63, depthkernelTI.f90:57-60)
! selects them with DAZI_DEPTHKERNEL_BACKEND=hip / DAZI_DEPTHKERNEL_TI_BACKEND=hip.
! Outputs are written in place in the reference's layouts; the tile controls
! DAZI_PDSURF_TILE_COLUMNS / DAZI_DEPTHKERNEL_TI_TILE_COLUMNS are read as the
! reference reads them (CalSurfG.f90:173-174, depthkernelTI.f90:250-253).
! A non-zero ABI code prints the library's message (stdout and unit 66) and
! stops with that code (9301-9304).  DAZI_DEPTHKERNEL_DUMP_DIR and
! DAZI_DEPTHKERNEL_TI_DUMP_DIR write the reference's dump files.
module dazi_dk
  use, intrinsic :: iso_c_binding
  implicit none
  private
  public :: depthkernel_hip, depthkernelTI_hip, dazi_dk_desc, dazi_dk_tune
  public :: c_dazi_dk_iso, c_dazi_dk_ti, c_dazi_dk_set_tune, dazi_dk_error_text

  integer(c_int32_t), parameter, public :: DAZI_DK_OK = 0
  integer(c_int32_t), parameter, public :: DAZI_DK_EINVAL = 9301
  integer(c_int32_t), parameter, public :: DAZI_DK_EHIP = 9302
  integer(c_int32_t), parameter, public :: DAZI_DK_ELIMIT = 9303
  integer(c_int32_t), parameter, public :: DAZI_DK_ENODEV = 9304

  type, bind(c) :: dazi_dk_desc
    integer(c_int32_t) :: nx, ny, nz, kmax, iwave, igr, tile_columns, vp_rho_mode
    real(c_float) :: minthk0
    integer(c_int32_t) :: reserved
  end type dazi_dk_desc

  type, bind(c) :: dazi_dk_tune
    integer(c_int32_t) :: lanes_per_scan, cols_per_wg, mode, tile_columns
    real(c_double) :: refine_tol
  end type dazi_dk_tune

  interface
    function c_dazi_dk_iso(desc, vel, vp, rho, depz, tRc, pvRc, sen_vs, sen_vp, sen_rho) &
        bind(c, name='dazi_dk_iso') result(rc)
      import :: c_int32_t, c_float, c_double, c_ptr, dazi_dk_desc
      type(dazi_dk_desc), intent(in) :: desc
      real(c_float), intent(in) :: vel(*)
      type(c_ptr), value :: vp, rho
      real(c_float), intent(in) :: depz(*)
      real(c_double), intent(in) :: tRc(*)
      real(c_double), intent(inout) :: pvRc(*), sen_vs(*), sen_vp(*), sen_rho(*)
      integer(c_int32_t) :: rc
    end function c_dazi_dk_iso

    function c_dazi_dk_ti(desc, vel, depz, tRc, pvRc, lsen) bind(c, name='dazi_dk_ti') result(rc)
      import :: c_int32_t, c_float, c_double, dazi_dk_desc
      type(dazi_dk_desc), intent(in) :: desc
      real(c_float), intent(in) :: vel(*), depz(*)
      real(c_double), intent(in) :: tRc(*)
      real(c_double), intent(inout) :: pvRc(*)
      real(c_float), intent(inout) :: lsen(*)
      integer(c_int32_t) :: rc
    end function c_dazi_dk_ti

    function c_dazi_dk_set_tune(tune) bind(c, name='dazi_dk_set_tune') result(rc)
      import :: c_int32_t, dazi_dk_tune
      type(dazi_dk_tune), intent(in) :: tune
      integer(c_int32_t) :: rc
    end function c_dazi_dk_set_tune

    function c_dazi_dk_last_error() bind(c, name='dazi_dk_last_error') result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function c_dazi_dk_last_error

    function c_strlen(s) bind(c, name='strlen') result(n)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: n
    end function c_strlen
  end interface

contains

  ! Text of the library's last error.
  function dazi_dk_error_text() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: n, i
    p = c_dazi_dk_last_error()
    if (.not. c_associated(p)) then
      msg = ''
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, s, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = s(i)
    end do
  end function dazi_dk_error_text

  integer function env_int(name, default)
    character(len=*), intent(in) :: name
    integer, intent(in) :: default
    character(len=64) :: v
    integer :: st, ios
    env_int = default
    call get_environment_variable(name, v, status=st)
    if (st == 0 .and. len_trim(v) > 0) then
      read(v, *, iostat=ios) env_int
      if (ios /= 0) env_int = default
    end if
    if (env_int < 1) env_int = 1
  end function env_int

  subroutine log_line(text)
    character(len=*), intent(in) :: text
    logical :: log_open
    write(*, '(A)') text
    inquire(unit=66, opened=log_open)
    if (log_open) write(66, '(A)') text
  end subroutine log_line

  ! Writes the error to stdout and unit 66 (as the reference's backends do) and
  ! stops with the ABI code.
  subroutine check(rc, where)
    integer(c_int32_t), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(len=:), allocatable :: msg
    character(len=16) :: code
    if (rc == DAZI_DK_OK) return
    write(code, '(I0)') rc
    msg = 'ERROR: ' // where // ' hip backend: dazi_dk error ' // trim(code) // ': ' // &
          dazi_dk_error_text()
    call log_line(msg)
    select case (rc)
    case (DAZI_DK_EINVAL)
      stop 9301
    case (DAZI_DK_EHIP)
      stop 9302
    case (DAZI_DK_ELIMIT)
      stop 9303
    case (DAZI_DK_ENODEV)
      stop 9304
    case default
      stop 9300
    end select
  end subroutine check

  ! Dump directory from `var` (set and non-empty), with the reference's log line.
  logical function dump_dir_of(var, label, dir)
    character(len=*), intent(in) :: var, label
    character(len=:), allocatable, intent(out) :: dir
    character(len=1024) :: v
    integer :: st
    logical :: log_open
    call get_environment_variable(var, v, status=st)
    dump_dir_of = (st == 0 .and. len_trim(v) > 0)
    dir = trim(adjustl(v))
    if (st /= 0) return
    inquire(unit=66, opened=log_open)
    if (log_open) write(66, '(A,A)') label, dir
    if (dump_dir_of) call execute_command_line('mkdir -p ' // dir)
  end function dump_dir_of

  subroutine dump_f64(path, a, n)
    character(len=*), intent(in) :: path
    integer, intent(in) :: n
    real*8, intent(in) :: a(n)
    integer :: u
    open(newunit=u, file=path, form='unformatted', access='stream', status='replace')
    write(u) a
    close(u)
  end subroutine dump_f64

  subroutine dump_f32(path, a, n)
    character(len=*), intent(in) :: path
    integer, intent(in) :: n
    real*4, intent(in) :: a(n)
    integer :: u
    open(newunit=u, file=path, form='unformatted', access='stream', status='replace')
    write(u) a
    close(u)
  end subroutine dump_f32

  ! Reference: depthkernel (CalSurfG.f90:1-113), all three node kernels; dumps
  ! as depthkernel_dump_arrays (CalSurfG.f90:115-140) under DAZI_DEPTHKERNEL_DUMP_DIR.
  subroutine depthkernel_hip(nx, ny, nz, vel, pvRc, sen_vsRc, sen_vpRc, sen_rhoRc, &
                             iwave, igr, kmaxRc, tRc, depz, minthk)
    integer nx, ny, nz
    real vel(nx, ny, nz)
    integer kmaxRc
    real*8 sen_vpRc(ny*nx, kmaxRc, nz), sen_vsRc(ny*nx, kmaxRc, nz), sen_rhoRc(ny*nx, kmaxRc, nz)
    integer iwave, igr
    real minthk
    real depz(nz)
    real*8 tRc(kmaxRc)
    real*8 pvRc(nx*ny, kmaxRc)
    type(dazi_dk_desc) :: d
    character(len=160) :: line
    character(len=:), allocatable :: dir

    d = dazi_dk_desc(nx, ny, nz, kmaxRc, iwave, igr, &
                     env_int('DAZI_PDSURF_TILE_COLUMNS', 64), 0, minthk, 0)
    call log_line('Depth kernel backend: hip')
    write(line, '(A,I0,A,I0,A,I0,A,I0)') 'Depth kernel grid/periods: nx=', nx, ' ny=', ny, &
        ' nz=', nz, ' kmax=', kmaxRc
    call log_line(trim(line))
    call check(c_dazi_dk_iso(d, vel, c_null_ptr, c_null_ptr, depz, tRc, pvRc, &
                             sen_vsRc, sen_vpRc, sen_rhoRc), 'depthkernel')
    if (dump_dir_of('DAZI_DEPTHKERNEL_DUMP_DIR', 'Depth kernel dump dir: ', dir)) then
      call dump_f64(dir // '/pv_depthkernel.bin', pvRc, nx*ny*kmaxRc)
      call dump_f64(dir // '/sen_vs_depthkernel.bin', sen_vsRc, nx*ny*kmaxRc*nz)
      call dump_f64(dir // '/sen_vp_depthkernel.bin', sen_vpRc, nx*ny*kmaxRc*nz)
      call dump_f64(dir // '/sen_rho_depthkernel.bin', sen_rhoRc, nx*ny*kmaxRc*nz)
    end if
  end subroutine depthkernel_hip

  ! Reference: depthkernelTI (depthkernelTI.f90:2-34): pvRc and Lsen_Gsc.
  subroutine depthkernelTI_hip(nx, ny, nz, vel, pvRc, iwave, igr, kmaxRc, tRc, depz, minthk, &
                               Lsen_Gsc)
    integer nx, ny, nz
    real vel(nx, ny, nz)
    integer iwave, igr
    real minthk
    real depz(nz)
    integer kmaxRc
    real*8 tRc(kmaxRc)
    real*8 pvRc(nx*ny, kmaxRc)
    real*4 Lsen_Gsc(nx*ny, kmaxRc, nz-1)
    type(dazi_dk_desc) :: d
    character(len=160) :: line
    character(len=:), allocatable :: dir

    d = dazi_dk_desc(nx, ny, nz, kmaxRc, iwave, igr, &
                     env_int('DAZI_DEPTHKERNEL_TI_TILE_COLUMNS', 64), 0, minthk, 0)
    call log_line('Depth kernel TI backend: hip')
    write(line, '(A,I0,A,I0,A,I0,A,I0)') 'Depth kernel TI grid/periods: nx=', nx, ' ny=', ny, &
        ' nz=', nz, ' kmax=', kmaxRc
    call log_line(trim(line))
    write(line, '(A,I0)') 'Depth kernel TI batch tile columns: ', d%tile_columns
    call log_line(trim(line))
    call check(c_dazi_dk_ti(d, vel, depz, tRc, pvRc, Lsen_Gsc), 'depthkernelTI')
    if (dump_dir_of('DAZI_DEPTHKERNEL_TI_DUMP_DIR', 'Depth kernel TI dump dir: ', dir)) then
      call dump_f64(dir // '/pv_ti.bin', pvRc, nx*ny*kmaxRc)
      call dump_f32(dir // '/lsen_gsc_ti.bin', Lsen_Gsc, nx*ny*kmaxRc*(nz-1))
    end if
  end subroutine depthkernelTI_hip

end module dazi_dk
