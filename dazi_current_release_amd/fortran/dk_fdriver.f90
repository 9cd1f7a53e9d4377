! dk_fdriver -- example Fortran host of the depth-kernel stage: reads a grid
! from a stream file, calls depthkernel_hip and depthkernelTI_hip (the
! reference's argument lists) and writes their outputs to a stream file.
!   input : nx, ny, nz, kmax (int32), minthk (f32), vel(nx,ny,nz) f32,
!           depz(nz) f32, tRc(kmax) f64
!   output: pvRc(nx*ny,kmax) f64, sen_vs, sen_vp, sen_rho (nx*ny,kmax,nz) f64,
!           pvTI(nx*ny,kmax) f64, Lsen_Gsc(nx*ny,kmax,nz-1) f32
! With two more arguments it then assembles G rows as the reference's G
! builder would, ray by ray through g_rows_hip_add_ray and one
! g_rows_hip_flush:
!   rays   : nray (int32), knumi(nray) int32 (1-based), then per ray the maps
!            fdm, fdmc, fdms (0:ny-1, 0:nx-1) f32 (the reference's
!            fdm(0:nvz+1, 0:nvx+1))
!   output : nar (int64), rw(nar) f32, iw(nar) int32, col(nar) int32
! (a rays argument "-" skips this step).  With two more arguments it calls
! FwdAzimuthalAniMap_hip on the TI output:
!   azm in : goxd, gozd, dvxd, dvzd (f32), gcf, gsf (nx-2,ny-2,nz-1) f32,
!            tRcV((nx-2)*(ny-2),kmax) f64
!   azm out: the unit-42 text lines
! (an azm argument "-" skips it).  With two more arguments it computes the
! receiver traveltimes on the depthkernel_hip phase velocities through
! surface_traveltimes_hip, with the G builder's geometry arrays:
!   trv in : goxd, gozd, dvxd, dvzd (f32), nsrcsurf, nrcf, kmaxd (int32),
!            nsrcsurf1(kmaxd), periods(nsrcsurf,kmaxd), nrc1(nsrcsurf,kmaxd)
!            int32, scxf, sczf (nsrcsurf,kmaxd) f32, rcxf, rczf
!            (nrcf,nsrcsurf,kmaxd) f32
!   trv out: ndata (int32), dsurf(ndata) f32
program dk_fdriver
  use dazi_dk
  implicit none
  integer :: nx, ny, nz, kmax, u
  real :: minthk
  real, allocatable :: vel(:,:,:), depz(:)
  real*8, allocatable :: tRc(:), pv(:,:), svs(:,:,:), svp(:,:,:), srho(:,:,:), pvti(:,:)
  real*4, allocatable :: lsen(:,:,:)
  character(len=1024) :: fin, fout, frays, fg
  integer :: nray, r
  integer, allocatable :: knumi(:), iw(:), col(:)
  real, allocatable :: fdm(:,:), fdmc(:,:), fdms(:,:), rw(:)
  integer(kind=8) :: nar, maxnar
  character(len=1024) :: fazin, fazout
  real :: goxd, gozd, dvxd, dvzd
  real, allocatable :: gcf(:,:,:), gsf(:,:,:)
  real*8, allocatable :: tRcV(:,:)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), form='unformatted', access='stream', status='old')
  read(u) nx, ny, nz, kmax, minthk
  allocate(vel(nx, ny, nz), depz(nz), tRc(kmax))
  read(u) vel, depz, tRc
  close(u)
  allocate(pv(nx*ny, kmax), svs(nx*ny, kmax, nz), svp(nx*ny, kmax, nz), srho(nx*ny, kmax, nz))
  allocate(pvti(nx*ny, kmax), lsen(nx*ny, kmax, nz-1))
  call depthkernel_hip(nx, ny, nz, vel, pv, svs, svp, srho, 2, 0, kmax, tRc, depz, minthk)
  call depthkernelTI_hip(nx, ny, nz, vel, pvti, 2, 0, kmax, tRc, depz, minthk, lsen)
  open(newunit=u, file=trim(fout), form='unformatted', access='stream', status='replace')
  write(u) pv, svs, svp, srho, pvti, lsen
  close(u)
  if (command_argument_count() < 4) stop
  call get_command_argument(3, frays)
  call get_command_argument(4, fg)
  if (trim(frays) /= '-') call g_rows_step()
  if (command_argument_count() < 6) stop
  call get_command_argument(5, fazin)
  call get_command_argument(6, fazout)
  if (trim(fazin) /= '-') then
    open(newunit=u, file=trim(fazin), form='unformatted', access='stream', status='old')
    allocate(gcf(nx-2, ny-2, nz-1), gsf(nx-2, ny-2, nz-1), tRcV((nx-2)*(ny-2), kmax))
    read(u) goxd, gozd, dvxd, dvzd, gcf, gsf, tRcV
    close(u)
    open(42, file=trim(fazout), status='replace', action='write')
    call FwdAzimuthalAniMap_hip(nx, ny, nz, 0, goxd, gozd, dvxd, dvzd, kmax, tRc, gcf, gsf, &
                                lsen, tRcV)
  end if
  if (command_argument_count() < 8) stop
  call travel_step()

contains

  subroutine travel_step()
    character(len=1024) :: ftin, ftout
    integer :: nsrcsurf, nrcf, kmaxd, ndata
    integer, allocatable :: nsrcsurf1(:), periods(:,:), nrc1(:,:)
    real, allocatable :: scxf(:,:), sczf(:,:), rcxf(:,:,:), rczf(:,:,:), dsurf(:)
    call get_command_argument(7, ftin)
    call get_command_argument(8, ftout)
    open(newunit=u, file=trim(ftin), form='unformatted', access='stream', status='old')
    read(u) goxd, gozd, dvxd, dvzd, nsrcsurf, nrcf, kmaxd
    allocate(nsrcsurf1(kmaxd), periods(nsrcsurf, kmaxd), nrc1(nsrcsurf, kmaxd))
    allocate(scxf(nsrcsurf, kmaxd), sczf(nsrcsurf, kmaxd), rcxf(nrcf, nsrcsurf, kmaxd), &
             rczf(nrcf, nsrcsurf, kmaxd), dsurf(nrcf * nsrcsurf * kmaxd))
    read(u) nsrcsurf1, periods, nrc1, scxf, sczf, rcxf, rczf
    close(u)
    call surface_traveltimes_hip(nx, ny, goxd, gozd, dvxd, dvzd, kmax, pv, nsrcsurf, nrcf, &
                                 kmaxd, nsrcsurf1, periods, nrc1, scxf, sczf, rcxf, rczf, &
                                 dsurf, ndata)
    open(newunit=u, file=trim(ftout), form='unformatted', access='stream', status='replace')
    write(u) ndata, dsurf(1:ndata)
    close(u)
  end subroutine travel_step

  subroutine g_rows_step()
  open(newunit=u, file=trim(frays), form='unformatted', access='stream', status='old')
  read(u) nray
  allocate(knumi(nray), fdm(0:ny-1, 0:nx-1), fdmc(0:ny-1, 0:nx-1), fdms(0:ny-1, 0:nx-1))
  read(u) knumi
  do r = 1, nray
    read(u) fdm, fdmc, fdms
    call g_rows_hip_add_ray(knumi(r), ny - 2, nx - 2, fdm, fdmc, fdms)
  end do
  close(u)
  maxnar = 3_8 * (nz - 1) * (nx - 2) * (ny - 2) * nray
  allocate(rw(maxnar), iw(maxnar), col(maxnar))
  nar = 0
  call g_rows_hip_flush(nx, ny, nz, kmax, vel, svs, svp, srho, lsen, 1, maxnar, rw, iw, col, nar)
  open(newunit=u, file=trim(fg), form='unformatted', access='stream', status='replace')
  write(u) nar, rw(1:nar), iw(1:nar), col(1:nar)
  close(u)
  end subroutine g_rows_step
end program dk_fdriver
