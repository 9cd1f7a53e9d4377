! dk_fdriver -- example Fortran host of the depth-kernel stage: reads a grid
! from a stream file, calls depthkernel_hip and depthkernelTI_hip (the
! reference's argument lists) and writes their outputs to a stream file.
!   input : nx, ny, nz, kmax (int32), minthk (f32), vel(nx,ny,nz) f32,
!           depz(nz) f32, tRc(kmax) f64
!   output: pvRc(nx*ny,kmax) f64, sen_vs, sen_vp, sen_rho (nx*ny,kmax,nz) f64,
!           pvTI(nx*ny,kmax) f64, Lsen_Gsc(nx*ny,kmax,nz-1) f32
program dk_fdriver
  use dazi_dk
  implicit none
  integer :: nx, ny, nz, kmax, u
  real :: minthk
  real, allocatable :: vel(:,:,:), depz(:)
  real*8, allocatable :: tRc(:), pv(:,:), svs(:,:,:), svp(:,:,:), srho(:,:,:), pvti(:,:)
  real*4, allocatable :: lsen(:,:,:)
  character(len=1024) :: fin, fout

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
end program dk_fdriver
