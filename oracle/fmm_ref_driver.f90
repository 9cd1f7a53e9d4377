! Test infrastructure only (linked into oracle/_ref/libdazi_ref.so).
!
! Drives the reference's own traveltime routines, compiled from
! code/src_inv_iso_joint/CalSurfG.f90 (module globalp :292-358, module traveltime
! with travel/fouds1/fouds2/addtree/downtree/updtree :375-1034, gridder
! :1596-1689, bsplrefine :1698-1770, srtimes :1772-1906), for ONE (period,
! source) pair, with the grid set-up of CalSurfGAniso_Joint.f90:375-432 and the
! per-source sequence of CalSurfGAniso_Joint.f90:513-699 (coarse gridding, the
! refined source grid, the refined and the coarse fast-marching passes) and the
! receiver times of :720-725.  The sequence is restated here; every numerical
! step is the reference's routine.
!
! Outputs: the coarse traveltime field ttn(nnz,nnx), the refined field and its
! node status as the coarse pass sees them (ttnr, nstsr over nrnz x nrnx, when
! asgr = 1), and one time per receiver.  stat = 0 on success.
subroutine ref_fmm_job(nvx_, nvz_, goxd_, gozd_, dvxd_, dvzd_, gdx_, gdz_, sgdl_, sgs_, &
                       asgr_, fom_, snb_, earth_, pv, scx, scz, nrc, rcx, rcz, &
                       ttn_out, nrnx_out, nrnz_out, ttnr_out, nstsr_out, times, stat) &
                       bind(C, name='ref_fmm_job')
  use iso_c_binding
  use globalp
  use traveltime
  implicit none
  integer(c_int), value :: nvx_, nvz_, gdx_, gdz_, sgdl_, sgs_, asgr_, fom_, nrc
  real(c_float), value :: goxd_, gozd_, dvxd_, dvzd_, snb_, earth_, scx, scz
  real(c_double), intent(in) :: pv(*)
  real(c_float), intent(in) :: rcx(*), rcz(*)
  real(c_float), intent(out) :: ttn_out(*), ttnr_out(*), times(*)
  integer(c_int), intent(out) :: nrnx_out, nrnz_out, nstsr_out(*), stat
  integer :: maxbt, urg, isx, isz, sw, sgs, nnxb, nnzb, idm1, idm2, ogx, ogz, grdfx, grdfz
  integer :: i, j, k, l
  real :: x, z, dnxb, dnzb, goxb, gozb, rx, rz, t1

  stat = 0
  gdx = gdx_
  gdz = gdz_
  asgr = asgr_
  sgdl = sgdl_
  sgs = sgs_
  earth = earth_
  fom = fom_
  snb = snb_
  goxd = goxd_
  gozd = gozd_
  dvxd = dvxd_
  dvzd = dvzd_
  nvx = nvx_
  nvz = nvz_
  if (allocated(velv)) deallocate(velv)
  if (allocated(veln)) deallocate(veln)
  if (allocated(ttn)) deallocate(ttn)
  if (allocated(nsts)) deallocate(nsts)
  if (allocated(btg)) deallocate(btg)
  if (allocated(velnb)) deallocate(velnb)
  if (allocated(ttnr)) deallocate(ttnr)
  if (allocated(nstsr)) deallocate(nstsr)
  allocate(velv(0:nvz+1, 0:nvx+1))
  dvx = dvxd*pi/180.0
  dvz = dvzd*pi/180.0
  gox = (90.0-goxd)*pi/180.0
  goz = gozd*pi/180.0
  nnx = (nvx-1)*gdx+1
  nnz = (nvz-1)*gdz+1
  dnx = dvx/gdx
  dnz = dvz/gdz
  allocate(veln(nnz, nnx), ttn(nnz, nnx), nsts(nnz, nnx))
  maxbt = nint(snb*nnx*nnz)
  allocate(btg(maxbt))

  call gridder(pv)
  x = scx
  z = scz
  nrnx_out = 0
  nrnz_out = 0
  if (asgr == 1) then
    allocate(velnb(nnz, nnx))
    velnb(1:nnz, 1:nnx) = veln(1:nnz, 1:nnx)
    nnxb = nnx
    nnzb = nnz
    dnxb = dnx
    dnzb = dnz
    goxb = gox
    gozb = goz
    isx = int((x-gox)/dnx)+1
    isz = int((z-goz)/dnz)+1
    if (isx < 1 .or. isx > nnx .or. isz < 1 .or. isz > nnz) then
      stat = 1
      return
    end if
    if (isx == nnx) isx = isx-1
    if (isz == nnz) isz = isz-1
    vnl = max(isx-sgs, 1)
    vnr = min(isx+sgs, nnx)
    vnt = max(isz-sgs, 1)
    vnb = min(isz+sgs, nnz)
    nrnx = (vnr-vnl)*sgdl+1
    nrnz = (vnb-vnt)*sgdl+1
    drnx = dvx/real(gdx*sgdl)
    drnz = dvz/real(gdz*sgdl)
    gorx = gox+dnx*(vnl-1)
    gorz = goz+dnz*(vnt-1)
    nnx = nrnx
    nnz = nrnz
    dnx = drnx
    dnz = drnz
    gox = gorx
    goz = gorz
    if (nnx > nnxb .or. nnz > nnzb) then
      idm1 = max(nnx, nnxb)
      idm2 = max(nnz, nnzb)
      deallocate(veln, ttn, nsts, btg)
      allocate(veln(idm2, idm1), ttn(idm2, idm1), nsts(idm2, idm1))
      maxbt = nint(snb*idm1*idm2)
      allocate(btg(maxbt))
    end if
    call bsplrefine
    urg = 1
    call travel(x, z, urg)
    allocate(ttnr(nnzb, nnxb), nstsr(nnzb, nnxb))
    if (nnx > nnxb .or. nnz > nnzb) then
      idm1 = max(nnx, nnxb)
      idm2 = max(nnz, nnzb)
      deallocate(ttnr, nstsr)
      allocate(ttnr(idm2, idm1), nstsr(idm2, idm1))
    end if
    ttnr = ttn
    nstsr = nsts
    nrnx_out = nnx
    nrnz_out = nnz
    do i = 1, nnx
      do j = 1, nnz
        ttnr_out(j+(i-1)*nnz) = ttnr(j, i)
        nstsr_out(j+(i-1)*nnz) = nstsr(j, i)
      end do
    end do
    ogx = vnl
    ogz = vnt
    grdfx = sgdl
    grdfz = sgdl
    nsts = -1
    do k = 1, nnz, grdfz
      idm1 = ogz+(k-1)/grdfz
      do l = 1, nnx, grdfx
        idm2 = ogx+(l-1)/grdfx
        nsts(idm1, idm2) = nstsr(k, l)
        if (nsts(idm1, idm2) >= 0) ttn(idm1, idm2) = ttnr(k, l)
      end do
    end do
    nnxr = nnx
    nnzr = nnz
    goxr = gox
    gozr = goz
    dnxr = dnx
    dnzr = dnz
    nnx = nnxb
    nnz = nnzb
    dnx = dnxb
    dnz = dnzb
    gox = goxb
    goz = gozb
    do j = 1, nnx
      do k = 1, nnz
        veln(k, j) = velnb(k, j)
      end do
    end do
    do k = 1, nnx
      do l = 1, nnz
        if (nsts(l, k) == 0) then
          if (l-1 >= 1) then
            if (nsts(l-1, k) == -1) nsts(l, k) = 1
          end if
          if (l+1 <= nnz) then
            if (nsts(l+1, k) == -1) nsts(l, k) = 1
          end if
          if (k-1 >= 1) then
            if (nsts(l, k-1) == -1) nsts(l, k) = 1
          end if
          if (k+1 <= nnx) then
            if (nsts(l, k+1) == -1) nsts(l, k) = 1
          end if
        end if
      end do
    end do
    urg = 2
    call travel(x, z, urg)
  else
    urg = 0
    call travel(x, z, urg)
  end if
  do i = 1, nnx
    do j = 1, nnz
      ttn_out(j+(i-1)*nnz) = ttn(j, i)
    end do
  end do
  do i = 1, nrc
    rx = rcx(i)
    rz = rcz(i)
    call srtimes(x, z, rx, rz, t1)
    times(i) = t1
  end do
end subroutine ref_fmm_job
