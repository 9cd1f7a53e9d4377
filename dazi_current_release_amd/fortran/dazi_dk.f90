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
!
! g_rows_hip_add_ray / g_rows_hip_flush replace the G builder's row fill and
! compression (CalSurfGAniso_Joint.f90:746-797): inside the record loop, in
! place of the row fill, add_ray takes the ray's fdm, fdmc, fdms maps and
! period index; flush assembles every buffered ray on the GPU and appends the
! (rw, iw, col) triplets at nar+1 in the reference's order.
! DAZI_G_TRIPLET_DUMP_DIR is left to the caller's dump_g_triplets.
!
! FwdAzimuthalAniMap_hip takes the argument list of FwdAzimuthalAniMap
! (FwdAzimuthalAniMap.f90:6-8) and writes the same unit-42 lines.
module dazi_dk
  use, intrinsic :: iso_c_binding
  implicit none
  private
  public :: depthkernel_hip, depthkernelTI_hip, dazi_dk_desc, dazi_dk_tune
  public :: g_rows_hip_add_ray, g_rows_hip_flush, dazi_dk_grow_desc, c_dazi_dk_grows
  public :: FwdAzimuthalAniMap_hip, dazi_dk_azim_desc, c_dazi_dk_azimuth_map
  public :: surface_traveltimes_hip, dazi_dk_travel_desc, c_dazi_dk_travel
  public :: c_dazi_dk_iso, c_dazi_dk_ti, c_dazi_dk_set_tune, dazi_dk_error_text

  integer(c_int32_t), parameter, public :: DAZI_DK_OK = 0
  integer(c_int32_t), parameter, public :: DAZI_DK_EINVAL = 9301
  integer(c_int32_t), parameter, public :: DAZI_DK_EHIP = 9302
  integer(c_int32_t), parameter, public :: DAZI_DK_ELIMIT = 9303
  integer(c_int32_t), parameter, public :: DAZI_DK_ENODEV = 9304

  type, bind(c) :: dazi_dk_desc
    integer(c_int32_t) :: nx, ny, nz, kmax, iwave, igr, tile_columns, vp_rho_mode
    real(c_float) :: minthk0
    integer(c_int32_t) :: flags
  end type dazi_dk_desc

  type, bind(c) :: dazi_dk_tune
    integer(c_int32_t) :: lanes_per_scan, cols_per_wg, mode, tile_columns
    real(c_double) :: refine_tol
  end type dazi_dk_tune

  type, bind(c) :: dazi_dk_grow_desc
    integer(c_int32_t) :: nvx, nvz, nz, kmax
    integer(c_int64_t) :: nray
    integer(c_int32_t) :: row_base
    real(c_float) :: ftol
  end type dazi_dk_grow_desc

  type, bind(c) :: dazi_dk_travel_desc
    integer(c_int32_t) :: nvx, nvz, gdx, gdz, sgdl, sgs, asgr, fom
    real(c_float) :: goxd, gozd, dvxd, dvzd, snb, earth
    integer(c_int32_t) :: kmax, slots
  end type dazi_dk_travel_desc

  type, bind(c) :: dazi_dk_azim_desc
    integer(c_int32_t) :: nx, ny, nz, kmax
  end type dazi_dk_azim_desc

  ! rays buffered by g_rows_hip_add_ray (cells in the reference's jj, kk order)
  integer :: gb_nray = 0, gb_ncell = 0, gb_nvx = 0, gb_nvz = 0
  integer(c_int32_t), allocatable :: gb_k(:), gb_cell(:)
  integer(c_int64_t), allocatable :: gb_off(:)
  real(c_float), allocatable :: gb_w(:), gb_wc(:), gb_ws(:)
  real, parameter :: GROW_FTOL = 1e-4   ! CalSurfGAniso_Joint.f90:330

  interface
    function c_dazi_dk_grows(desc, ray_k, cell_off, cell, fdm, fdmc, fdms, vel, sen_vs, sen_vp, &
                             sen_rho, lsen, cap, rw, iw, col, nar) &
        bind(c, name='dazi_dk_grows') result(rc)
      import :: c_int32_t, c_int64_t, c_float, c_double, dazi_dk_grow_desc
      type(dazi_dk_grow_desc), intent(in) :: desc
      integer(c_int32_t), intent(in) :: ray_k(*), cell(*)
      integer(c_int64_t), intent(in) :: cell_off(*)
      real(c_float), intent(in) :: fdm(*), fdmc(*), fdms(*), vel(*), lsen(*)
      real(c_double), intent(in) :: sen_vs(*), sen_vp(*), sen_rho(*)
      integer(c_int64_t), value :: cap
      real(c_float), intent(inout) :: rw(*)
      integer(c_int32_t), intent(inout) :: iw(*), col(*)
      integer(c_int64_t), intent(out) :: nar
      integer(c_int32_t) :: rc
    end function c_dazi_dk_grows

    function c_dazi_dk_azimuth_map(desc, lsen, gcf, gsf, tRcV, cos_rc, sin_rc, amp, amp_rela, &
                                   angle) bind(c, name='dazi_dk_azimuth_map') result(rc)
      import :: c_int32_t, c_float, c_double, dazi_dk_azim_desc
      type(dazi_dk_azim_desc), intent(in) :: desc
      real(c_float), intent(in) :: lsen(*), gcf(*), gsf(*)
      real(c_double), intent(in) :: tRcV(*)
      real(c_float), intent(out) :: cos_rc(*), sin_rc(*), amp(*), amp_rela(*), angle(*)
      integer(c_int32_t) :: rc
    end function c_dazi_dk_azimuth_map

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

    function c_dazi_dk_travel(desc, njob, job_k, scx, scz, rec_off, rcx, rcz, pvRc, times, &
                              ttn, status) bind(c, name='dazi_dk_travel') result(rc)
      import :: c_int32_t, c_int64_t, c_float, c_double, c_ptr, dazi_dk_travel_desc
      type(dazi_dk_travel_desc), intent(in) :: desc
      integer(c_int64_t), value :: njob
      integer(c_int32_t), intent(in) :: job_k(*)
      real(c_float), intent(in) :: scx(*), scz(*), rcx(*), rcz(*)
      integer(c_int64_t), intent(in) :: rec_off(*)
      real(c_double), intent(in) :: pvRc(*)
      real(c_float), intent(out) :: times(*)
      type(c_ptr), value :: ttn
      integer(c_int32_t), intent(out) :: status(*)
      integer(c_int32_t) :: rc
    end function c_dazi_dk_travel

    function c_dazi_dk_enable_timing(on) bind(c, name='dazi_dk_enable_timing') result(rc)
      import :: c_int32_t
      integer(c_int32_t), value :: on
      integer(c_int32_t) :: rc
    end function c_dazi_dk_enable_timing

    function c_dazi_dk_last_timings(out, n) bind(c, name='dazi_dk_last_timings') result(m)
      import :: c_int32_t, c_double
      real(c_double), intent(out) :: out(*)
      integer(c_int32_t), value :: n
      integer(c_int32_t) :: m
    end function c_dazi_dk_last_timings

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

  ! True when environment variable `name` is set to a non-empty value other than 0.
  logical function env_on(name)
    character(len=*), intent(in) :: name
    character(len=64) :: v
    integer :: st
    call get_environment_variable(name, v, status=st)
    env_on = st == 0 .and. len_trim(v) > 0 .and. trim(adjustl(v)) /= '0'
  end function env_on

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
    real(c_double) :: tm(7)
    integer(c_int32_t) :: rc, nt, lref
    integer(c_int32_t), parameter :: TI_FAILED_NAN = 2   ! DAZI_DK_TI_FAILED_NAN

    d = dazi_dk_desc(nx, ny, nz, kmaxRc, iwave, igr, &
                     env_int('DAZI_DEPTHKERNEL_TI_TILE_COLUMNS', 64), 0, minthk, 0)
    ! the reference leaves NaN in Lsen_Gsc at periods without a phase velocity
    ! (tregn96 at c = 0); on request this backend does the same, else 0
    if (env_on('DAZI_DEPTHKERNEL_TI_FAILED_NAN')) d%flags = ior(d%flags, TI_FAILED_NAN)
    call log_line('Depth kernel TI backend: hip')
    write(line, '(A,I0,A,I0,A,I0,A,I0)') 'Depth kernel TI grid/periods: nx=', nx, ' ny=', ny, &
        ' nz=', nz, ' kmax=', kmaxRc
    call log_line(trim(line))
    write(line, '(A,I0)') 'Depth kernel TI batch tile columns: ', d%tile_columns
    call log_line(trim(line))
    rc = c_dazi_dk_enable_timing(1_c_int32_t)
    rc = c_dazi_dk_ti(d, vel, depz, tRc, pvRc, Lsen_Gsc)
    nt = 0
    if (rc == DAZI_DK_OK) nt = c_dazi_dk_last_timings(tm, 7_c_int32_t)
    lref = c_dazi_dk_enable_timing(0_c_int32_t)
    call check(rc, 'depthkernelTI')
    ! the reference's internal timing lines (depthkernelTI.f90:170-186), from
    ! the library's event timings: model prep and root search, the TI kernels
    ! (tregn96 and the splice are one kernel here), the refined layer count
    ! (every column of a node model has the same) and the TI launch's threads
    if (nt >= 7) then
      lref = nint(tm(6))
      call log_line('  depthkernelTI internal timing:')
      write(line, '(a,f12.3,a)') '    refine+surfdisp96 phase velocity = ', &
          (tm(1) + tm(2)) * 1.0d-3, ' s'
      call log_line(trim(line))
      write(line, '(a,f12.3,a)') '    tregn96 sensitivity kernel = ', tm(4) * 1.0d-3, ' s'
      call log_line(trim(line))
      write(line, '(a,f12.3,a)') '    splice refined kernel = ', 0.0d0, ' s'
      call log_line(trim(line))
      write(line, '(a,i0,a,i0,a,f10.2)') '    refined sublayers rmax min/max/avg = ', lref, '/', &
          lref, '/', dble(lref)
      call log_line(trim(line))
      write(line, '(a,i0)') '    observed worker threads = ', nint(tm(7))
      call log_line(trim(line))
    end if
    if (dump_dir_of('DAZI_DEPTHKERNEL_TI_DUMP_DIR', 'Depth kernel TI dump dir: ', dir)) then
      call dump_f64(dir // '/pv_ti.bin', pvRc, nx*ny*kmaxRc)
      call dump_f32(dir // '/lsen_gsc_ti.bin', Lsen_Gsc, nx*ny*kmaxRc*(nz-1))
    end if
  end subroutine depthkernelTI_hip

  ! Buffers one ray of period index knumi (1-based, as the reference's knumi):
  ! the cells (jj, kk) with |fdm(jj,kk)| >= ftol, in the reference's loop order
  ! (jj outer, kk inner), with their fdm, fdmc, fdms.
  subroutine g_rows_hip_add_ray(knumi, nvz, nvx, fdm, fdmc, fdms)
    integer, intent(in) :: knumi, nvz, nvx
    real, intent(in) :: fdm(0:nvz+1, 0:nvx+1), fdmc(0:nvz+1, 0:nvx+1), fdms(0:nvz+1, 0:nvx+1)
    integer :: jj, kk, n
    if (gb_nray == 0) then
      gb_nvx = nvx
      gb_nvz = nvz
      if (.not. allocated(gb_k)) then
        allocate(gb_k(1024), gb_off(1025), gb_cell(65536), gb_w(65536), gb_wc(65536), &
                 gb_ws(65536))
      end if
      gb_off(1) = 0
      gb_ncell = 0
    else if (nvx /= gb_nvx .or. nvz /= gb_nvz) then
      write(*, '(A)') 'g_rows_hip_add_ray: the grid changed within one batch'
      stop DAZI_DK_EINVAL
    end if
    if (gb_nray + 1 > size(gb_k)) then
      call grow_i32(gb_k, 2 * size(gb_k))
      call grow_i64(gb_off, 2 * size(gb_off))
    end if
    n = count(abs(fdm(1:nvz, 1:nvx)) >= GROW_FTOL)
    if (gb_ncell + n > size(gb_cell)) then
      call grow_i32(gb_cell, max(2 * size(gb_cell), gb_ncell + n))
      call grow_f32(gb_w, size(gb_cell))
      call grow_f32(gb_wc, size(gb_cell))
      call grow_f32(gb_ws, size(gb_cell))
    end if
    do jj = 1, nvz
      do kk = 1, nvx
        if (abs(fdm(jj, kk)) >= GROW_FTOL) then
          gb_ncell = gb_ncell + 1
          gb_cell(gb_ncell) = (jj - 1) * nvx + (kk - 1)
          gb_w(gb_ncell) = fdm(jj, kk)
          gb_wc(gb_ncell) = fdmc(jj, kk)
          gb_ws(gb_ncell) = fdms(jj, kk)
        end if
      end do
    end do
    gb_nray = gb_nray + 1
    gb_k(gb_nray) = knumi - 1
    gb_off(gb_nray + 1) = gb_ncell
  end subroutine g_rows_hip_add_ray

  ! Assembles the buffered rays (data rows first_row, first_row+1, ...) from the
  ! stage outputs in the reference's layouts and appends their triplets to
  ! rw/iw/col at nar+1 (at most maxnar entries in all); nar is updated and the
  ! buffer emptied.
  subroutine g_rows_hip_flush(nx, ny, nz, kmaxRc, vels, sen_vsRc, sen_vpRc, sen_rhoRc, Lsen_Gsc, &
                              first_row, maxnar, rw, iw, col, nar)
    integer, intent(in) :: nx, ny, nz, kmaxRc, first_row
    real, intent(in) :: vels(nx, ny, nz)
    real*8, intent(in) :: sen_vsRc(nx*ny, kmaxRc, nz), sen_vpRc(nx*ny, kmaxRc, nz)
    real*8, intent(in) :: sen_rhoRc(nx*ny, kmaxRc, nz)
    real*4, intent(in) :: Lsen_Gsc(nx*ny, kmaxRc, nz-1)
    integer(kind=8), intent(in) :: maxnar
    real, intent(inout) :: rw(*)
    integer, intent(inout) :: iw(*), col(*)
    integer(kind=8), intent(inout) :: nar
    type(dazi_dk_grow_desc) :: d
    integer(c_int64_t) :: n
    integer(c_int32_t) :: dummy_i(1)
    real(c_float) :: dummy_f(1)
    if (gb_nray == 0) return
    if (gb_nvx /= nx - 2 .or. gb_nvz /= ny - 2) then
      write(*, '(A)') 'g_rows_hip_flush: the rays do not match the nx x ny grid'
      stop DAZI_DK_EINVAL
    end if
    call log_line('G row assembly backend: hip')
    d = dazi_dk_grow_desc(nx - 2, ny - 2, nz, kmaxRc, int(gb_nray, c_int64_t), first_row, &
                          GROW_FTOL)
    if (maxnar > nar) then
      call check(c_dazi_dk_grows(d, gb_k, gb_off, gb_cell, gb_w, gb_wc, gb_ws, vels, sen_vsRc, &
                                 sen_vpRc, sen_rhoRc, Lsen_Gsc, maxnar - nar, rw(nar + 1), &
                                 iw(nar + 1), col(nar + 1), n), 'g_rows')
    else
      call check(c_dazi_dk_grows(d, gb_k, gb_off, gb_cell, gb_w, gb_wc, gb_ws, vels, sen_vsRc, &
                                 sen_vpRc, sen_rhoRc, Lsen_Gsc, 0_c_int64_t, dummy_f, dummy_i, &
                                 dummy_i, n), 'g_rows')
    end if
    nar = nar + n
    gb_nray = 0
  end subroutine g_rows_hip_flush

  ! Reference: FwdAzimuthalAniMap (FwdAzimuthalAniMap.f90:6-89): the map on the
  ! GPU, then the same lines on unit 42 (opened by the caller), which is closed.
  ! Replaces the traveltime part of the joint G builder's period/source/record
  ! loops (CalSurfGAniso_Joint.f90:511-725: gridder, the refined source grid,
  ! travel, srtimes) with one call for every (period, source) job, in the loops'
  ! order: knumi = 1..kmax, srcnum = 1..nsrcsurf1(knumi), istep =
  ! 1..nrc1(srcnum, knumi).  dsurf(count1) receives the receiver times in that
  ! order (ndata of them), bit-identical to the reference's srtimes.  The grid
  ! constants are the builder's (gdx = gdz = 5, sgdl = sgs = 8, asgr = 1,
  ! fom = 1, snb = 0.5, earth = 6371).  A source or receiver outside the model
  ! stops as the reference does; a job whose reference result is undefined
  ! (status 2, 3 or 4 of dazi_dk_travel) stops with its index.
  subroutine surface_traveltimes_hip(nx, ny, goxd, gozd, dvxd, dvzd, kmaxRc, pvRc, nsrcsurf, &
                                     nrcf, kmax, nsrcsurf1, periods, nrc1, scxf, sczf, rcxf, &
                                     rczf, dsurf, ndata)
    integer nx, ny, kmaxRc, nsrcsurf, nrcf, kmax, ndata
    real goxd, gozd, dvxd, dvzd
    real*8 pvRc(nx*ny, kmaxRc)
    integer nsrcsurf1(kmax), periods(nsrcsurf, kmax), nrc1(nsrcsurf, kmax)
    real scxf(nsrcsurf, kmax), sczf(nsrcsurf, kmax)
    real rcxf(nrcf, nsrcsurf, kmax), rczf(nrcf, nsrcsurf, kmax)
    real dsurf(*)
    type(dazi_dk_travel_desc) :: d
    integer(c_int32_t), allocatable :: job_k(:), st(:)
    real(c_float), allocatable :: scx(:), scz(:), rcx(:), rcz(:), times(:)
    integer(c_int64_t), allocatable :: off(:)
    integer :: knumi, srcnum, istep, q, njob, nrec
    njob = 0
    nrec = 0
    do knumi = 1, kmax
      do srcnum = 1, nsrcsurf1(knumi)
        njob = njob + 1
        nrec = nrec + nrc1(srcnum, knumi)
      end do
    end do
    allocate(job_k(max(njob, 1)), st(max(njob, 1)), scx(max(njob, 1)), scz(max(njob, 1)), &
             off(njob + 1), rcx(max(nrec, 1)), rcz(max(nrec, 1)), times(max(nrec, 1)))
    q = 0
    off(1) = 0
    do knumi = 1, kmax
      do srcnum = 1, nsrcsurf1(knumi)
        q = q + 1
        job_k(q) = periods(srcnum, knumi) - 1
        scx(q) = scxf(srcnum, knumi)
        scz(q) = sczf(srcnum, knumi)
        do istep = 1, nrc1(srcnum, knumi)
          rcx(off(q) + istep) = rcxf(istep, srcnum, knumi)
          rcz(off(q) + istep) = rczf(istep, srcnum, knumi)
        end do
        off(q + 1) = off(q) + nrc1(srcnum, knumi)
      end do
    end do
    d = dazi_dk_travel_desc(nx - 2, ny - 2, 5, 5, 8, 8, 1, 1, goxd, gozd, dvxd, dvzd, 0.5, &
                            6371.0, kmaxRc, 0)
    call check(c_dazi_dk_travel(d, int(njob, c_int64_t), job_k, scx, scz, off, rcx, rcz, pvRc, &
                                times, c_null_ptr, st), 'surface traveltimes')
    do q = 1, njob
      if (st(q) /= 0) then
        write(6, *) 'surface traveltimes: job ', q, ' has status ', st(q), &
                    ' (the reference result is undefined there)'
        stop 9303
      end if
    end do
    dsurf(1:nrec) = times(1:nrec)
    ndata = nrec
  end subroutine surface_traveltimes_hip

  subroutine FwdAzimuthalAniMap_hip(nx, ny, nz, maxvp, goxd, gozd, dvxd, dvzd, kmaxRc, tRc, &
                                    gcf, gsf, Lsen_Gsc, tRcV)
    integer nx, ny, nz
    integer maxvp
    real goxd, gozd, dvxd, dvzd
    integer kmaxRc
    real*8 tRc(kmaxRc)
    real gcf(nx-2, ny-2, nz-1), gsf(nx-2, ny-2, nz-1)
    real*8 tRcV((nx-2)*(ny-2), kmaxRc)
    real*4 Lsen_Gsc(nx*ny, kmaxRc, nz-1)
    real, allocatable :: cosRc(:,:), sinRc(:,:), amp(:,:), rela(:,:), ang(:,:)
    type(dazi_dk_azim_desc) :: d
    integer :: ii, jj, tt, c
    real :: isoC
    allocate(cosRc((nx-2)*(ny-2), kmaxRc), sinRc((nx-2)*(ny-2), kmaxRc), &
             amp((nx-2)*(ny-2), kmaxRc), rela((nx-2)*(ny-2), kmaxRc), ang((nx-2)*(ny-2), kmaxRc))
    d = dazi_dk_azim_desc(nx, ny, nz, kmaxRc)
    call check(c_dazi_dk_azimuth_map(d, Lsen_Gsc, gcf, gsf, tRcV, cosRc, sinRc, amp, rela, &
                                     ang), 'FwdAzimuthalAniMap')
    do tt = 1, kmaxRc
      do jj = 1, ny-2
        do ii = 1, nx-2
          c = (jj-1)*(nx-2) + ii
          isoC = tRcV(c, tt)
          write(42, '(10f10.5)') gozd+(jj-1)*dvzd, goxd-(ii-1)*dvxd, tRc(tt), isoC, &
              ang(c, tt), rela(c, tt), amp(c, tt), cosRc(c, tt), sinRc(c, tt)
        end do
      end do
    end do
    close(42)
  end subroutine FwdAzimuthalAniMap_hip

  subroutine grow_i32(a, n)
    integer(c_int32_t), allocatable, intent(inout) :: a(:)
    integer, intent(in) :: n
    integer(c_int32_t), allocatable :: b(:)
    allocate(b(n))
    b(1:size(a)) = a
    call move_alloc(b, a)
  end subroutine grow_i32

  subroutine grow_i64(a, n)
    integer(c_int64_t), allocatable, intent(inout) :: a(:)
    integer, intent(in) :: n
    integer(c_int64_t), allocatable :: b(:)
    allocate(b(n))
    b(1:size(a)) = a
    call move_alloc(b, a)
  end subroutine grow_i64

  subroutine grow_f32(a, n)
    real(c_float), allocatable, intent(inout) :: a(:)
    integer, intent(in) :: n
    real(c_float), allocatable :: b(:)
    allocate(b(n))
    b(1:size(a)) = a
    call move_alloc(b, a)
  end subroutine grow_f32

end module dazi_dk
