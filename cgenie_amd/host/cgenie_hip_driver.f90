! cgenie_hip_driver.f90 -- runs the tracer path in the genie.f90 order
! (genie.f90:306-313 GOLDSTEIN every ocean step; 360-381 BIOGEM and the
! tracer coupling every kbiogem ocean steps) through the cgenie_hip module,
! with the argument lists of the reference's wrappers
! (genie_loop_wrappers.f90:143-150, 313-315, 321): step_goldstein with zero
! surface heat and freshwater fluxes and the given (or zero) wind stress,
! step_biogem with the BIOGEM step in seconds and the interface arrays of an
! ocean-only run, biogem_tracercoupling.
!
!   cgenie_hip_driver INPUT OUTPUT
!
! INPUT is a little-endian stream: int32 maxi, maxj, maxk, maxl, nsteps,
! kbiogem, diso, precision, has_bgc, device, mirror, has_mom, has_sfx; float64 dphi, rdphi, dsc,
! diff(2), ec(5); int32 k1(0:maxi+1,0:maxj+1); float64 rc, rc2, cv, cv2, rds,
! dsv, rdsv, ds, sv (each 0:maxj+1); float64 dt, rdz, rdza, dza, ssmax, dz
! (each 0:maxk+1); float64 go_u(3,0:maxi,0:maxj,maxk),
! go_ts(maxl,maxi,maxj,maxk); if has_bgc: int32 slots(6) (1-based, 0 =
! absent), float64 dtyr, carb_tol, light(maxi,maxj), mld(maxi,maxj); if
! has_mom (the velocity solve): float64 s, c, rcv, rds2 (each 0:maxj+1),
! int32 labels(maxi,0:maxj), isles, npi(isles), paths(3,sum npi), float64
! adrag, rel, scf, stress(maxi,maxj,4) (stressxu, stressyu, stressxv, stressyv);
! if has_sfx (EMBM's surface fluxes, surflux_hip before every step_goldstein
! as genie.f90:271-282 calls surflux): int32 nyear, float64
! solfor(maxj,nyear), albcl, ca, pmeadj (maxi,maxj), int32 iroff, jroff
! (maxi,maxj), igrid, polar_avg, radfor_atchem, float64 scl(17) (the order
! of initialise_surflux_hip), atm(maxi,maxj,5) (atemp, ashum, sich, sica,
! tice), and, without has_mom, stress(maxi,maxj,4).
! OUTPUT receives float64 go_ts(maxl,maxi,maxj,maxk), inventory(maxl),
! go_rho(maxi,maxj,maxk), the energy and water diagnostics of the last
! step_goldstein, dum_sfcocn1(1:38,maxi,maxj) and dum_sfxsed1(1:34,maxi,maxj)
! of the last step_biogem (both zero without mirroring), then go_u(3,maxi,
! maxj,maxk), go_psi(0:maxi,0:maxj) and go_tau(2,maxi,maxj) of the last
! step_goldstein; with has_sfx then the last surflux call's outputs
! (maxi,maxj,28) in the order of CGO_SF_* and its ashum and tice (maxi,maxj).
program cgenie_hip_driver
  use, intrinsic :: iso_c_binding
  use cgenie_hip
  implicit none
  integer(c_int32_t) :: hdr(13)
  integer :: maxi, maxj, maxk, maxl, nsteps, kbiogem, device, koverall, u
  logical :: diso, has_bgc, has_mom, has_sfx
  integer(c_int32_t) :: nyear, sfx_i(3)
  integer(c_int32_t), allocatable :: rout(:,:,:)
  real(c_double), allocatable :: solfor(:,:), stab(:,:,:), atm(:,:,:), sfa(:,:,:), sfxo(:,:,:)
  real(c_double) :: scl(17)
  real(c_double), allocatable :: mt(:,:), stress(:,:,:)
  integer(c_int32_t), allocatable :: labels(:,:), npi(:), paths(:,:)
  integer(c_int32_t) :: isles
  real(c_double) :: msc(3)
  real(c_double) :: sc(10), dtyr, carb_tol
  integer(c_int32_t), allocatable :: k1(:,:)
  real(c_double), allocatable :: jt(:,:), kt(:,:), go_u(:,:,:,:), go_ts(:,:,:,:), go_ts1(:,:,:,:)
  real(c_double), allocatable :: light(:,:), mld(:,:), inv(:)
  integer(c_int32_t) :: slots(6)
  ! the genie_global arrays of the calls (genie_global.f90): surface fluxes
  ! and outputs (ilon1_ocn, ilat1_ocn), go_uvw, go_cost, go_tau, go_psi,
  ! go_mldta, go_rho, and the BIOGEM interface arrays (tracers, lon, lat)
  real(c_double), allocatable :: f2(:,:,:), go_uvw(:,:,:,:), go_cost(:,:), go_tau(:,:,:), go_psi(:,:)
  real(c_double), allocatable :: go_mldta(:,:), go_rho(:,:,:), sfcatm(:,:,:), sfxatm(:,:,:), sfcocn(:,:,:)
  real(c_double), allocatable :: sfxocn(:,:,:), sfcsed(:,:,:), sfxsed(:,:,:), sfxrok(:,:,:)
  real(c_double) :: test_energy, test_water, dts
  integer(c_int64_t) :: clock
  real(c_double), parameter :: conv_yr_s = 3600.0_c_double * (24.0_c_double * 365.25_c_double)
  character(len=4096) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
  read (u) hdr
  maxi = hdr(1); maxj = hdr(2); maxk = hdr(3); maxl = hdr(4); nsteps = hdr(5); kbiogem = max(1, int(hdr(6)))
  diso = hdr(7) /= 0; has_bgc = hdr(9) /= 0; device = hdr(10); has_mom = hdr(12) /= 0; has_sfx = hdr(13) /= 0
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
  if (has_mom) then
     allocate (mt(0:maxj+1, 4), labels(maxi, 0:maxj), stress(maxi, maxj, 4))
     read (u) mt
     read (u) labels
     read (u) isles
     allocate (npi(max(isles, 1)))
     npi = 0
     if (isles > 0) read (u) npi(1:isles)
     allocate (paths(3, max(sum(npi), 1)))
     paths = 0
     if (isles > 0) read (u) paths(:, 1:sum(npi))
     read (u) msc
     read (u) stress
  end if
  if (has_sfx) then
     read (u) nyear
     allocate (solfor(maxj, nyear), stab(maxi, maxj, 3), rout(maxi, maxj, 2), atm(maxi, maxj, 5))
     read (u) solfor
     read (u) stab
     read (u) rout
     read (u) sfx_i
     read (u) scl
     read (u) atm
     if (.not. has_mom) then
        allocate (stress(maxi, maxj, 4))
        read (u) stress
     end if
  end if
  close (u)
  go_ts1 = go_ts

  call initialise_goldstein_hip(device, maxi, maxj, maxk, maxl, int(k1), sc(1), sc(2), &
       jt(:, 1), jt(:, 2), jt(:, 3), jt(:, 4), jt(:, 5), jt(:, 6), jt(:, 7), jt(:, 8), jt(:, 9), &
       kt(:, 1), kt(:, 2), kt(:, 3), kt(:, 4), kt(:, 5), kt(:, 6), sc(6:10), sc(3), sc(4:5), diso, &
       int(hdr(8)), go_ts, go_u)
  if (has_mom) call initialise_goldstein_momentum_hip(mt(:, 1), mt(:, 2), mt(:, 3), mt(:, 4), int(labels), &
       int(isles), int(npi), int(paths(:, 1:max(sum(npi), 0))), msc(1), msc(2), msc(3))
  if (has_sfx) call initialise_surflux_hip(solfor, stab(:, :, 1), stab(:, :, 2), stab(:, :, 3), int(rout(:, :, 1)), &
       int(rout(:, :, 2)), int(sfx_i(1)), int(sfx_i(2)), sfx_i(3) /= 0, scl)
  if (has_bgc) call initialise_biogem_hip(int(slots), dtyr, light, mld, carb_tol)
  call cgenie_hip_set_mirror(hdr(11) /= 0)

  allocate (f2(maxi, maxj, 18), go_uvw(3, maxi, maxj, maxk), go_cost(maxi, maxj), go_tau(2, maxi, maxj))
  allocate (go_psi(0:maxi, 0:maxj), go_mldta(maxi, maxj), go_rho(maxi, maxj, maxk))
  allocate (sfcatm(19, maxi, maxj), sfxatm(19, maxi, maxj), sfcocn(95, maxi, maxj), sfxocn(95, maxi, maxj))
  allocate (sfcsed(79, maxi, maxj), sfxsed(79, maxi, maxj), sfxrok(95, maxi, maxj))
  f2 = 0; go_cost = 0; go_tau = 0; go_psi = 0; go_mldta = 0; go_rho = 0
  sfcatm = 0; sfxatm = 0; sfcocn = 0; sfxocn = 0; sfcsed = 0; sfxsed = 0; sfxrok = 0
  go_uvw = go_u(:, 1:maxi, 1:maxj, :)
  if (has_mom .or. has_sfx) f2(:, :, 10:13) = stress
  ! surflux's atmosphere-side outputs: runoff_land, fxlha, fxsha, fxswa,
  ! fxlwa, evap_atm, pptn_atm, dthsic, dtareasic, albice
  allocate (sfa(maxi, maxj, 10), sfxo(maxi, maxj, 28))
  sfa = 0
  dts = dtyr * conv_yr_s    ! dtyr is the BIOGEM step (kbiogem ocean steps)
  clock = 0
  do koverall = 1, nsteps
     if (has_sfx) call surflux_hip(koverall, f2(:,:,14), f2(:,:,15), atm(:,:,1), atm(:,:,2), atm(:,:,3), &
          atm(:,:,4), atm(:,:,5), sfa(:,:,10), f2(:,:,10), f2(:,:,11), f2(:,:,12), f2(:,:,13), f2(:,:,18), &
          f2(:,:,1), f2(:,:,2), f2(:,:,3), f2(:,:,4), f2(:,:,6), f2(:,:,7), f2(:,:,8), sfa(:,:,1), &
          sfa(:,:,2), sfa(:,:,3), sfa(:,:,4), sfa(:,:,5), sfa(:,:,6), sfa(:,:,7), sfa(:,:,8), sfa(:,:,9))
     call step_goldstein(koverall, f2(:,:,1), f2(:,:,2), f2(:,:,3), f2(:,:,4), f2(:,:,5), f2(:,:,6), &
          f2(:,:,7), f2(:,:,8), f2(:,:,9), f2(:,:,10), f2(:,:,11), f2(:,:,12), f2(:,:,13), &
          f2(:,:,14), f2(:,:,15), f2(:,:,16), f2(:,:,17), f2(:,:,18), test_energy, test_water, &
          go_ts, go_ts1, go_cost, go_uvw, go_tau, go_psi, go_mldta, go_rho)
     if (has_bgc .and. mod(koverall, kbiogem) == 0) then
        call step_biogem(dts, clock, sfcatm, sfxatm, sfcocn, sfxocn, sfcsed, sfxsed, sfxrok)
        call biogem_tracercoupling(go_ts, go_ts1)
     end if
  end do

  call cgenie_hip_download(go_ts, go_ts1)
  call cgenie_hip_inventory(inv)
  if (has_sfx) call cgenie_hip_surflux_outputs(sfxo)
  call end_cgenie_hip()
  open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
  write (u) go_ts
  write (u) inv
  write (u) go_rho
  write (u) test_energy, test_water
  write (u) sfcocn(1:38, :, :)
  write (u) sfxsed(1:34, :, :)
  write (u) go_uvw
  write (u) go_psi
  write (u) go_tau
  if (has_sfx) then
     write (u) sfxo
     write (u) atm(:, :, 2)
     write (u) atm(:, :, 5)
  end if
  close (u)
  write (*, '(a,i0,a)') 'cgenie_hip_driver: ', nsteps, ' ocean steps done'
end program cgenie_hip_driver
