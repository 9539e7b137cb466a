! cgenie_hip_driver.f90 -- runs the tracer path in the genie.f90 order
! (genie.f90:306-313 GOLDSTEIN every ocean step; 360-381 BIOGEM and the
! tracer coupling every kbiogem ocean steps) through the cgenie_hip module.
!
!   cgenie_hip_driver INPUT OUTPUT
!
! INPUT is a little-endian stream: int32 maxi, maxj, maxk, maxl, nsteps,
! kbiogem, diso, precision, has_bgc, device, mirror; float64 dphi, rdphi, dsc,
! diff(2), ec(5); int32 k1(0:maxi+1,0:maxj+1); float64 rc, rc2, cv, cv2, rds,
! dsv, rdsv, ds, sv (each 0:maxj+1); float64 dt, rdz, rdza, dza, ssmax, dz
! (each 0:maxk+1); float64 go_u(3,0:maxi,0:maxj,maxk),
! go_ts(maxl,maxi,maxj,maxk); if has_bgc: int32 slots(6) (1-based, 0 =
! absent), float64 dtyr, carb_tol, light(maxi,maxj), mld(maxi,maxj).
! OUTPUT receives float64 go_ts(maxl,maxi,maxj,maxk) then inventory(maxl).
program cgenie_hip_driver
  use, intrinsic :: iso_c_binding
  use cgenie_hip
  implicit none
  integer(c_int32_t) :: hdr(11)
  integer :: maxi, maxj, maxk, maxl, nsteps, kbiogem, device, koverall, u
  logical :: diso, has_bgc
  real(c_double) :: sc(10), dtyr, carb_tol
  integer(c_int32_t), allocatable :: k1(:,:)
  real(c_double), allocatable :: jt(:,:), kt(:,:), go_u(:,:,:,:), go_ts(:,:,:,:), go_ts1(:,:,:,:)
  real(c_double), allocatable :: light(:,:), mld(:,:), inv(:)
  integer(c_int32_t) :: slots(6)
  character(len=4096) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
  read (u) hdr
  maxi = hdr(1); maxj = hdr(2); maxk = hdr(3); maxl = hdr(4); nsteps = hdr(5); kbiogem = max(1, int(hdr(6)))
  diso = hdr(7) /= 0; has_bgc = hdr(9) /= 0; device = hdr(10)
  read (u) sc
  allocate (k1(0:maxi+1, 0:maxj+1), jt(0:maxj+1, 9), kt(0:maxk+1, 6))
  allocate (go_u(3, 0:maxi, 0:maxj, maxk), go_ts(maxl, maxi, maxj, maxk), go_ts1(maxl, maxi, maxj, maxk))
  allocate (light(maxi, maxj), mld(maxi, maxj), inv(maxl))
  read (u) k1
  read (u) jt
  read (u) kt
  read (u) go_u
  read (u) go_ts
  if (has_bgc) then
     read (u) slots
     read (u) dtyr, carb_tol
     read (u) light
     read (u) mld
  end if
  close (u)
  go_ts1 = go_ts

  call initialise_goldstein_hip(device, maxi, maxj, maxk, maxl, int(k1), sc(1), sc(2), &
       jt(:, 1), jt(:, 2), jt(:, 3), jt(:, 4), jt(:, 5), jt(:, 6), jt(:, 7), jt(:, 8), jt(:, 9), &
       kt(:, 1), kt(:, 2), kt(:, 3), kt(:, 4), kt(:, 5), kt(:, 6), sc(6:10), sc(3), sc(4:5), diso, &
       int(hdr(8)), go_ts, go_u)
  if (has_bgc) call initialise_biogem_hip(int(slots), dtyr, light, mld, carb_tol)
  call cgenie_hip_set_mirror(hdr(11) /= 0)

  do koverall = 1, nsteps
     call step_goldstein_tracers(go_ts, go_ts1)
     if (has_bgc .and. mod(koverall, kbiogem) == 0) then
        call step_biogem(dtyr)
        call biogem_tracercoupling(go_ts, go_ts1)
     end if
  end do

  call cgenie_hip_download(go_ts, go_ts1)
  call cgenie_hip_inventory(inv)
  call end_cgenie_hip()
  open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
  write (u) go_ts
  write (u) inv
  close (u)
  write (*, '(a,i0,a)') 'cgenie_hip_driver: ', nsteps, ' ocean steps done'
end program cgenie_hip_driver
