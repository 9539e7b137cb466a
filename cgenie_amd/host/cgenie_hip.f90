! cgenie_hip.f90 -- drop-in Fortran host module for the tracer path of one
! cGENIE ocean step on the MI355X engine.
!
! It keeps the reference's calling pattern: module state set up once (as
! initialise_goldstein / initialise_biogem do, goldstein.f90:906-1075,
! biogem_data.f90:1098-1140), then per ocean step the calls that genie.f90
! makes (genie.f90:306-313 GOLDSTEIN, 360-381 BIOGEM), with the reference's
! argument lists, so the wrappers' CALL lines (genie_loop_wrappers.f90:122-151,
! 310-322) compile unchanged against this module:
!
!   step_goldstein(istep, latent_ocn, ..., go_rho)   goldstein.f90:17-38: the
!                                         surface-flux slot from the heat and
!                                         freshwater fluxes (149-170), go_ts in,
!                                         with initialise_goldstein_momentum_hip
!                                         the velocity solve from the wind
!                                         stresses (102-107, 198-233), tstepo
!                                         (235), go_ts/go_rho/go_cost (and
!                                         go_u/go_tau/go_psi) and the surface
!                                         and energy/water outputs back (427-478)
!   step_biogem(dum_dts, dum_genie_clock, dum_sfcatm1, ..., dum_sfxsumrok1)
!                                         biogem.f90:528-547 (biology,
!                                         remineralisation and surface chemistry
!                                         of the carried species; bottom-water
!                                         composition and rain out, 1719-1754)
!   surflux_hip(istep, otemp, osaln, atemp, ashum, ..., dthsic, dtareasic)
!                                         surflux, embm.f90:2548-3738, without
!                                         ENTS (initialise_surflux_hip first):
!                                         the fluxes from the device's own top
!                                         level, the surface slot formed on the
!                                         device; step_goldstein then adds only
!                                         the sea-ice model's fluxes
!   biogem_tracercoupling(dum_ts, dum_ts1) biogem.f90:1885-1890
!   tstepo()                              goldstein.f90:2280-2432, called at 235
!   step_goldstein_tracers(go_ts, go_ts1) tstepo and the go_ts hand-back alone
!
! The tracer state lives on the GPU.  With host mirroring on (the default,
! the reference's behaviour) go_ts / go_ts1 are refreshed after every call
! that changes the state; with it off they are refreshed by
! cgenie_hip_download only, which is what a batched host wants.  Errors stop
! the program with the engine's message, as the reference's die does
! (genie_util.f90:15-32).
module cgenie_hip
  use, intrinsic :: iso_c_binding
  use cgo_api
  implicit none
  private

  public :: initialise_goldstein_hip, initialise_biogem_hip, end_cgenie_hip
  public :: tstepo, step_goldstein_tracers, step_biogem, biogem_tracercoupling
  public :: cgenie_hip_set_mirror, cgenie_hip_download, cgenie_hip_inventory, cgenie_hip_diag
  public :: cgenie_hip_set_sflux, step_goldstein, cgenie_hip_atmosphere, initialise_goldstein_momentum_hip
  public :: initialise_surflux_hip, surflux_hip, cgenie_hip_surflux_outputs

  type(c_ptr), save :: eng = c_null_ptr
  integer(c_int32_t), save :: grid_id = -1, mem(1) = -1
  integer, save :: nx = 0, ny = 0, nz = 0, nl = 0
  logical, save :: mirror = .true.
  logical, save :: airsea = .false.   ! CO2 and O2 exchange with the device atmosphere
  ! arrays whose addresses the grid descriptor holds during cgo_grid_create
  integer(c_int32_t), allocatable, target, save :: k1_c(:,:)
  real(c_double), allocatable, target, save :: jtab(:,:), ktab(:,:)
  ! step_goldstein: the velocities last handed to the engine, the initial
  ! energy and water inventories (goldstein.f90:79-96), the surface-flux slot
  real(c_double), allocatable, target, save :: u_last(:,:,:,:), sflux_c(:,:,:)
  real(c_double), save :: dphi_s = 0, dsc_s = 0, ini_energy = 0, ini_water = 0
  logical, save :: sflux_on = .false.
  ! the velocity solve: on, scf (ini_gold_nml), dzz (goldstein.f90:1035) and
  ! the wind stresses last handed to the engine
  logical, save :: mom_on = .false.
  real(c_double), save :: scf_s = 2.0_c_double, dzz_s = 0
  real(c_double), allocatable, save :: stress_last(:,:,:)
  ! the surface fluxes: on, and the host arrays their calls hand over
  logical, save :: sfx_on = .false.
  real(c_double), allocatable, target, save :: sfx_in(:,:,:), sfx_out(:,:,:), sic_c(:,:,:)
  ! step_biogem: go_ts positions of DIC, ALK, PO4, O2, DOM_C, DOM_P and the step length
  integer, save :: bgc_slots(6) = 0
  real(c_double), save :: bgc_dtyr = 0

  ! the reference's REAL is 8 bytes (-fdefault-real-8, which also makes its
  ! literals double): every literal here carries the kind
  integer, parameter :: dp = c_double
  ! GOLDSTEIN scales and conversions (goldstein_lib.f90:48-63, goldstein-defaults.nml:27, 60)
  real(dp), parameter :: usc = 0.05_dp, rsc = 6.37E6_dp, dsc_p = 5.0E3_dp, rh0sc = 1.0E3_dp, cpsc = 3981.1_dp
  real(dp), parameter :: saln0 = 34.9_dp, m2mm = 1000.0_dp, mm2m = 1.0_dp / m2mm, albocn = 0.05_dp
  real(dp), parameter :: rfluxsc = rsc / (dsc_p * usc * rh0sc * cpsc)
  real(dp), parameter :: fsc = 2 * 7.2921E-5_dp    ! goldstein_lib.f90:49
  ! BIOGEM constants (gem_cmn.f90:510-513, 688, 696, 719) and tracer indices (72-129, 190-255)
  real(dp), parameter :: conv_yr_s = 3600.0_dp * (24.0_dp * 365.25_dp), const_pi = 3.141592653589793_dp
  real(dp), parameter :: conv_s_yr = 1.0_dp / (3600.0_dp * (24.0_dp * 365.25_dp))
  integer, parameter :: ia_pCO2 = 3, ia_pO2 = 6    ! gem_cmn.f90 atmospheric tracer indices
  real(dp), parameter :: const_rEarth = 6.37E+06_dp, nullsmall = 0.999999E-19_dp
  integer, parameter :: io_of_slot(6) = [3, 12, 8, 10, 15, 20]    ! DIC, ALK, PO4, O2, DOM_C, DOM_P
  integer, parameter :: io_Ca = 35, io_B = 36, io_F = 37, io_SO4 = 38
  integer, parameter :: is_of(5) = [3, 8, 14, 33, 34]              ! POC, POP, CaCO3, POC_frac2, CaCO3_frac2

contains

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char) :: buf(512)
    character(len=512) :: msg
    integer :: i, ierr
    if (rc == CGO_OK) return
    msg = ' '
    if (c_associated(eng)) then
       ierr = cgo_last_error(eng, buf, int(512, c_size_t))
       do i = 1, 512
          if (buf(i) == c_null_char) exit
          msg(i:i) = buf(i)
       end do
    end if
    write (*, '(a,a,a,i0,a,a)') 'cgenie_hip: ', what, ' failed (', rc, '): ', trim(msg)
    error stop 1
  end subroutine check

  ! Engine, grid and one member.  Per-j arrays are (0:maxj+1), per-k arrays
  ! (0:maxk+1), k1 is (0:maxi+1,0:maxj+1), go_ts(maxl,maxi,maxj,maxk),
  ! go_u(3,0:maxi,0:maxj,maxk); diff(1:2) non-dimensional (goldstein.f90:1406-1407).
  subroutine initialise_goldstein_hip(device, maxi, maxj, maxk, maxl, k1, dphi, rdphi, rc, rc2, cv, cv2, rds, dsv, &
       rdsv, ds, sv, dt, rdz, rdza, dza, ssmax, dz, ec, dsc, diff, diso, precision, go_ts, go_u)
    integer, intent(in) :: device, maxi, maxj, maxk, maxl
    integer, intent(in) :: k1(0:maxi+1, 0:maxj+1)
    real(c_double), intent(in) :: dphi, rdphi, dsc
    real(c_double), intent(in), dimension(0:maxj+1) :: rc, rc2, cv, cv2, rds, dsv, rdsv, ds, sv
    real(c_double), intent(in), dimension(0:maxk+1) :: dt, rdz, rdza, dza, ssmax, dz
    real(c_double), intent(in) :: ec(5), diff(2)
    logical, intent(in) :: diso
    integer, intent(in) :: precision
    real(c_double), intent(in) :: go_ts(maxl, maxi, maxj, maxk)
    real(c_double), intent(in) :: go_u(3, 0:maxi, 0:maxj, maxk)
    type(cgo_grid_desc) :: g
    type(cgo_member_params) :: p

    nx = maxi; ny = maxj; nz = maxk; nl = maxl
    call check(cgo_create(int(device, c_int32_t), eng), 'cgo_create')
    allocate (k1_c(0:maxi+1, 0:maxj+1), jtab(0:maxj+1, 9), ktab(0:maxk+1, 6))
    k1_c = int(k1, c_int32_t)
    jtab(:, 1) = rc;  jtab(:, 2) = rc2; jtab(:, 3) = cv;   jtab(:, 4) = cv2; jtab(:, 5) = rds
    jtab(:, 6) = dsv; jtab(:, 7) = rdsv; jtab(:, 8) = ds;  jtab(:, 9) = sv
    ktab(:, 1) = dt;  ktab(:, 2) = rdz; ktab(:, 3) = rdza; ktab(:, 4) = dza; ktab(:, 5) = ssmax
    ktab(:, 6) = dz
    g%maxi = int(maxi, c_int32_t); g%maxj = int(maxj, c_int32_t); g%maxk = int(maxk, c_int32_t)
    g%k1 = c_loc(k1_c)
    g%dphi = dphi; g%rdphi = rdphi
    g%rc = c_loc(jtab(0, 1));  g%rc2 = c_loc(jtab(0, 2)); g%cv = c_loc(jtab(0, 3))
    g%cv2 = c_loc(jtab(0, 4)); g%rds = c_loc(jtab(0, 5)); g%dsv = c_loc(jtab(0, 6))
    g%rdsv = c_loc(jtab(0, 7)); g%ds = c_loc(jtab(0, 8)); g%sv = c_loc(jtab(0, 9))
    g%dt = c_loc(ktab(0, 1));  g%rdz = c_loc(ktab(0, 2)); g%rdza = c_loc(ktab(0, 3))
    g%dza = c_loc(ktab(0, 4)); g%ssmax = c_loc(ktab(0, 5)); g%dz = c_loc(ktab(0, 6))
    g%ec = ec
    g%dsc = dsc
    call check(cgo_grid_create(eng, g, grid_id), 'cgo_grid_create')
    p%diff = diff
    p%diso = merge(1_c_int32_t, 0_c_int32_t, diso)
    p%carry = 1
    call check(cgo_member_create(eng, grid_id, int(maxl, c_int32_t), int(precision, c_int32_t), p, mem(1)), &
         'cgo_member_create')
    call check(cgo_set_velocity(eng, mem(1), go_u), 'cgo_set_velocity')
    call check(cgo_upload_state(eng, mem(1), go_ts), 'cgo_upload_state')
    dphi_s = dphi; dsc_s = dsc
    allocate (u_last(3, 0:maxi, 0:maxj, maxk), sflux_c(maxl, maxi, maxj))
    u_last = go_u
    sflux_c = 0
    call energy_water(go_ts, ini_energy, ini_water)
  end subroutine initialise_goldstein_hip

  ! GOLDSTEIN's velocity solve for the member (goldstein.f90:198-233), set up
  ! as initialise_goldstein does (1359-1575, 1788-1814): the per-j metrics
  ! s, c, rcv, rds2 (0:maxj+1), the world's .psiles labels as
  ! labels(maxi,0:maxj) (gbold(i+j*maxi), 1505-1517), isles, the .paths
  ! lengths npi(isles) and points paths(3,sum npi) (1538-1574), and adrag
  ! (days), rel and scf of ini_gold_nml.  step_goldstein then solves the
  ! velocities from its wind stresses and hands back go_u, go_tau and go_psi
  ! as the reference does (102-107, 444-450).
  subroutine initialise_goldstein_momentum_hip(s, c, rcv, rds2, labels, isles, npi, paths, adrag, rel, scf)
    real(c_double), intent(in), dimension(0:ny+1) :: s, c, rcv, rds2
    integer, intent(in) :: labels(nx, 0:ny), isles
    integer, intent(in) :: npi(:), paths(:,:)
    real(c_double), intent(in) :: adrag, rel, scf
    type(cgo_momentum_desc) :: d
    real(c_double), allocatable, target :: js(:,:)
    integer(c_int32_t), allocatable, target :: lab(:,:), np(:), pth(:,:)

    if (.not. c_associated(eng)) call check(CGO_E_STATE, 'initialise_goldstein_momentum_hip: initialise_goldstein_hip first')
    allocate (js(0:ny+1, 4), lab(nx, 0:ny), np(max(isles, 1)), pth(3, max(size(paths, 2), 1)))
    js(:, 1) = s; js(:, 2) = c; js(:, 3) = rcv; js(:, 4) = rds2
    lab = int(labels, c_int32_t)
    np = 0
    if (isles > 0) np(1:isles) = int(npi(1:isles), c_int32_t)
    pth = 0
    if (size(paths, 2) > 0) pth(:, 1:size(paths, 2)) = int(paths, c_int32_t)
    d%s = c_loc(js(0, 1)); d%c = c_loc(js(0, 2)); d%rcv = c_loc(js(0, 3)); d%rds2 = c_loc(js(0, 4))
    d%labels = c_loc(lab)
    d%isles = int(isles, c_int32_t)
    d%npi = c_loc(np)
    d%paths = c_loc(pth)
    d%adrag = adrag
    d%rel = rel
    call check(cgo_grid_set_momentum(eng, grid_id, d), 'cgo_grid_set_momentum')
    mom_on = .true.
    scf_s = scf
    dzz_s = ktab(nz, 6) * ktab(nz - 1, 4) / 2     ! dz(maxk) dza(maxk-1) / 2
    if (allocated(stress_last)) deallocate (stress_last)
  end subroutine initialise_goldstein_momentum_hip

  ! EMBM's surface fluxes without ENTS (initialise_embm, embm.f90:198-1530):
  ! its tables solfor(maxj,nyear), albcl, ca, pmeadj, iroff, jroff
  ! (maxi,maxj), and the namelist constants in the order of
  ! cgo_surflux_desc: scl(17) = scf, rmax, delf2x, rdtdim, rsictscsf,
  ! olr_adj, olr_adj0, t_eqm, par_albsic_min, par_albsic_max, par_sich_max,
  ! rate_co2, rate_ch4, rate_n2o and the initial co2, ch4, n2o.  surflux_hip
  ! then computes the fluxes and the surface slot on the device.
  subroutine initialise_surflux_hip(solfor, albcl, ca, pmeadj, iroff, jroff, igrid, polar_avg, radfor_atchem, scl)
    real(c_double), intent(in), target :: solfor(:,:), albcl(nx, ny), ca(nx, ny), pmeadj(nx, ny)
    integer, intent(in) :: iroff(nx, ny), jroff(nx, ny), igrid, polar_avg
    logical, intent(in) :: radfor_atchem
    real(c_double), intent(in) :: scl(17)
    type(cgo_surflux_desc) :: d
    real(c_double), allocatable, target :: tab(:,:,:), sol(:,:)
    integer(c_int32_t), allocatable, target :: ro(:,:,:)

    if (.not. c_associated(eng)) call check(CGO_E_STATE, 'initialise_surflux_hip: initialise_goldstein_hip first')
    if (size(solfor, 1) /= ny) call check(CGO_E_ARG, 'initialise_surflux_hip: solfor(maxj, nyear) expected')
    allocate (sol(ny, size(solfor, 2)), tab(nx, ny, 3), ro(nx, ny, 2))
    sol = solfor
    tab(:, :, 1) = albcl; tab(:, :, 2) = ca; tab(:, :, 3) = pmeadj
    ro(:, :, 1) = int(iroff, c_int32_t); ro(:, :, 2) = int(jroff, c_int32_t)
    d%solfor = c_loc(sol)
    d%nyear = int(size(solfor, 2), c_int32_t)
    d%albcl = c_loc(tab(1, 1, 1)); d%ca = c_loc(tab(1, 1, 2)); d%pmeadj = c_loc(tab(1, 1, 3))
    d%iroff = c_loc(ro(1, 1, 1)); d%jroff = c_loc(ro(1, 1, 2))
    d%igrid = int(igrid, c_int32_t)
    d%polar_avg = int(polar_avg, c_int32_t)
    d%radfor_atchem = merge(1_c_int32_t, 0_c_int32_t, radfor_atchem)
    d%scf = scl(1); d%rmax = scl(2); d%delf2x = scl(3); d%rdtdim = scl(4); d%rsictscsf = scl(5)
    d%olr_adj = scl(6); d%olr_adj0 = scl(7); d%t_eqm = scl(8); d%albsic_min = scl(9); d%albsic_max = scl(10)
    d%sich_max = scl(11); d%rate_co2 = scl(12); d%rate_ch4 = scl(13); d%rate_n2o = scl(14)
    d%co2 = scl(15); d%ch4 = scl(16); d%n2o = scl(17)
    call check(cgo_grid_set_surflux(eng, grid_id, d), 'cgo_grid_set_surflux')
    if (.not. mom_on) dzz_s = ktab(nz, 6) * ktab(nz - 1, 4) / 2
    allocate (sfx_in(nx, ny, 5), sfx_out(nx, ny, CGO_SF_N), sic_c(nx, ny, 2))
    sfx_on = .true.
  end subroutine initialise_surflux_hip

  ! the wind stresses, handed to the engine when they differ from the last
  ! (the velocity solve and the surface fluxes' wind speed read them)
  subroutine set_stress(stressxu_ocn, stressyu_ocn, stressxv_ocn, stressyv_ocn)
    real(c_double), dimension(:,:), intent(in) :: stressxu_ocn, stressyu_ocn, stressxv_ocn, stressyv_ocn
    if (.not. allocated(stress_last)) then
       allocate (stress_last(nx, ny, 4))
       stress_last = huge(1.0_dp)
    end if
    if (any(stressxu_ocn /= stress_last(:, :, 1)) .or. any(stressyu_ocn /= stress_last(:, :, 2)) .or. &
        any(stressxv_ocn /= stress_last(:, :, 3)) .or. any(stressyv_ocn /= stress_last(:, :, 4))) then
       call check(cgo_set_wind(eng, mem(1), stressxu_ocn, stressyu_ocn, stressxv_ocn, stressyv_ocn, scf_s), &
            'cgo_set_wind')
       stress_last(:, :, 1) = stressxu_ocn; stress_last(:, :, 2) = stressyu_ocn
       stress_last(:, :, 3) = stressxv_ocn; stress_last(:, :, 4) = stressyv_ocn
    end if
  end subroutine set_stress

  ! surflux (embm.f90:2548-2578) with the arguments of its configuration
  ! without ENTS, in the reference's order.  otemp and osaln, the ocean's top
  ! level, are read on the device (the top level its last tstepo left); the
  ! air temperature and humidity and the sea-ice height, cover and surface
  ! temperature are the host's (EMBM and the sea-ice model step them), ashum
  ! and tice come back as the call leaves them.  The surface slot for the
  ! next step_goldstein is formed on the device.
  subroutine surflux_hip(istep, otemp, osaln, atemp, ashum, sich, sica, tice, albice, stressxu_ocn, stressyu_ocn, &
       stressxv_ocn, stressyv_ocn, albedo, fxlho, fxsho, fxswo, fxlwo, evap_ocn, pptn_ocn, runoff_ocn, runoff_land, &
       fxlha, fxsha, fxswa, fxlwa, evap_atm, pptn_atm, dthsic, dtareasic)
    integer, intent(in) :: istep
    real(c_double), dimension(:,:), intent(in) :: otemp, osaln, atemp, sich, sica
    real(c_double), dimension(:,:), intent(in) :: stressxu_ocn, stressyu_ocn, stressxv_ocn, stressyv_ocn
    real(c_double), dimension(:,:), intent(inout) :: ashum, tice
    real(c_double), dimension(:,:), intent(out) :: albice, albedo, fxlho, fxsho, fxswo, fxlwo, evap_ocn, &
         pptn_ocn, runoff_ocn, runoff_land, fxlha, fxsha, fxswa, fxlwa, evap_atm, pptn_atm, dthsic, dtareasic
    integer(c_int32_t) :: ids(1)

    if (.not. sfx_on) call check(CGO_E_STATE, 'surflux_hip: initialise_surflux_hip first')
    if (istep < 0 .or. size(otemp) /= nx * ny .or. size(osaln) /= nx * ny) &
         call check(CGO_E_ARG, 'surflux_hip: bad istep or field shape')
    call set_stress(stressxu_ocn, stressyu_ocn, stressxv_ocn, stressyv_ocn)
    sfx_in(:, :, 1) = atemp; sfx_in(:, :, 2) = ashum; sfx_in(:, :, 3) = sich
    sfx_in(:, :, 4) = sica; sfx_in(:, :, 5) = tice
    call check(cgo_set_surface(eng, mem(1), c_loc(sfx_in(1, 1, 1)), c_loc(sfx_in(1, 1, 2)), c_loc(sfx_in(1, 1, 3)), &
         c_loc(sfx_in(1, 1, 4)), c_loc(sfx_in(1, 1, 5)), c_null_ptr, c_null_ptr), 'cgo_set_surface')
    ids(1) = mem(1)
    call check(cgo_surflux(eng, ids, 1_c_int32_t), 'cgo_surflux')
    call check(cgo_download_surflux(eng, mem(1), c_loc(sfx_out), c_loc(sfx_in(1, 1, 2)), c_loc(sfx_in(1, 1, 5)), &
         c_null_ptr), 'cgo_download_surflux')
    ashum = sfx_in(:, :, 2)
    tice = sfx_in(:, :, 5)
    fxlho = sfx_out(:, :, 1 + CGO_SF_FXLHO); fxsho = sfx_out(:, :, 1 + CGO_SF_FXSHO)
    fxswo = sfx_out(:, :, 1 + CGO_SF_FXSWO); fxlwo = sfx_out(:, :, 1 + CGO_SF_FXLWO)
    evap_ocn = sfx_out(:, :, 1 + CGO_SF_EVAP_OCN); pptn_ocn = sfx_out(:, :, 1 + CGO_SF_PPTN_OCN)
    runoff_ocn = sfx_out(:, :, 1 + CGO_SF_RUNOFF_OCN); runoff_land = sfx_out(:, :, 1 + CGO_SF_RUNOFF_LAND)
    fxlha = sfx_out(:, :, 1 + CGO_SF_FXLHA); fxsha = sfx_out(:, :, 1 + CGO_SF_FXSHA)
    fxswa = sfx_out(:, :, 1 + CGO_SF_FXSWA); fxlwa = sfx_out(:, :, 1 + CGO_SF_FXLWA)
    evap_atm = sfx_out(:, :, 1 + CGO_SF_EVAP_ATM); pptn_atm = sfx_out(:, :, 1 + CGO_SF_PPTN_ATM)
    dthsic = sfx_out(:, :, 1 + CGO_SF_DTHSIC); dtareasic = sfx_out(:, :, 1 + CGO_SF_DTAREASIC)
    albedo = sfx_out(:, :, 1 + CGO_SF_ALBEDO); albice = sfx_out(:, :, 1 + CGO_SF_ALBICE)
  end subroutine surflux_hip

  ! the last surflux_hip call's outputs, (maxi, maxj, CGO_SF_N) in the order of CGO_SF_*
  subroutine cgenie_hip_surflux_outputs(out)
    real(c_double), intent(out) :: out(nx, ny, CGO_SF_N)
    out = sfx_out
  end subroutine cgenie_hip_surflux_outputs

  ! The total heat and water diagnostics of step_goldstein
  ! (goldstein.f90:79-96, 458-471), summed in the reference's order.
  subroutine energy_water(ts, energy, water)
    real(c_double), intent(in) :: ts(nl, nx, ny, nz)
    real(c_double), intent(out) :: energy, water
    real(c_double) :: vsc
    integer :: j, k
    vsc = dphi_s * rsc * rsc
    energy = 0
    water = 0
    do k = 1, nz
       do j = 1, ny
          energy = energy + sum(ts(1, 1:nx, j, k)) * ktab(k, 6) * jtab(j, 8)
          water = water - sum(ts(2, 1:nx, j, k)) * ktab(k, 6) * jtab(j, 8)
       end do
    end do
    energy = energy * vsc * dsc_s * rh0sc * cpsc
    water = m2mm * water * vsc * dsc_s / saln0
  end subroutine energy_water

  ! BIOGEM state of the member from its current go_ts.  slots(1:6) are the
  ! 1-based go_ts positions of DIC, ALK, PO4, O2, DOM_C, DOM_P (0 = absent);
  ! light = solfor/solar_constant and mld (m) are (maxi,maxj).
  ! With wspeed (m/s, par_phys_windspeed) the carried CO2 and O2 exchange
  ! with the well-mixed atmosphere of ATCHEM on the device, starting from
  ! atm_init = (pCO2, pO2) in atm (atchem-defaults.nml) and stepped by every
  ! step_biogem; gas_a is par_gastransfer_a (biogem-defaults.nml:221).
  subroutine initialise_biogem_hip(slots, dtyr, light, mld, carb_tol, wspeed, atm_init, gas_a)
    integer, intent(in) :: slots(6)
    real(c_double), intent(in) :: dtyr
    real(c_double), intent(in) :: light(nx, ny), mld(nx, ny)
    real(c_double), intent(in), optional :: carb_tol
    real(c_double), intent(in), optional :: wspeed(nx, ny), atm_init(2), gas_a
    type(cgo_bgc_params) :: bp
    call cgo_bgc_default_params(bp, dtyr)
    bp%slot_DIC = int(slots(1) - 1, c_int32_t)
    bp%slot_ALK = int(slots(2) - 1, c_int32_t)
    bp%slot_PO4 = int(slots(3) - 1, c_int32_t)
    bp%slot_O2 = int(slots(4) - 1, c_int32_t)
    bp%slot_DOM_C = int(slots(5) - 1, c_int32_t)
    bp%slot_DOM_P = int(slots(6) - 1, c_int32_t)
    bp%caco3 = merge(1_c_int32_t, 0_c_int32_t, slots(1) > 0 .and. slots(2) > 0)
    if (present(carb_tol)) bp%carb_tol = carb_tol
    if (present(wspeed)) then
       bp%airsea = 1
       if (present(atm_init)) then
          bp%atm_pCO2 = atm_init(1)
          bp%atm_pO2 = atm_init(2)
       end if
       if (present(gas_a)) bp%gastransfer_a = gas_a
    end if
    call check(cgo_bgc_init(eng, mem(1), bp, light, mld), 'cgo_bgc_init')
    if (present(wspeed)) call check(cgo_set_windspeed(eng, mem(1), wspeed), 'cgo_set_windspeed')
    airsea = present(wspeed)
    bgc_slots = slots
    bgc_dtyr = dtyr
  end subroutine initialise_biogem_hip

  ! the surface boundary slot ts(:,1:maxi,1:maxj,maxk+1) (goldstein.f90:167-170)
  subroutine cgenie_hip_set_sflux(sflux)
    real(c_double), intent(in), target :: sflux(nl, nx, ny)
    call check(cgo_set_sflux(eng, mem(1), c_loc(sflux)), 'cgo_set_sflux')
  end subroutine cgenie_hip_set_sflux

  ! the device atmosphere: atm(1,:,:) pCO2, atm(2,:,:) pO2 in atm (ATCHEM's atm)
  subroutine cgenie_hip_atmosphere(atm)
    real(c_double), intent(out) :: atm(2, nx, ny)
    call check(cgo_bgc_download(eng, mem(1), 6_c_int32_t, atm), 'cgo_bgc_download atmosphere')
  end subroutine cgenie_hip_atmosphere

  subroutine cgenie_hip_set_mirror(on)
    logical, intent(in) :: on
    mirror = on
  end subroutine cgenie_hip_set_mirror

  subroutine cgenie_hip_download(go_ts, go_ts1)
    real(c_double), intent(inout) :: go_ts(nl, nx, ny, nz)
    real(c_double), intent(inout), optional :: go_ts1(nl, nx, ny, nz)
    call check(cgo_download_state(eng, mem(1), go_ts), 'cgo_download_state')
    if (present(go_ts1)) go_ts1 = go_ts
  end subroutine cgenie_hip_download

  ! per-tracer volume-weighted inventory (the diag2 weight, goldstein_diag.f90:39-41)
  subroutine cgenie_hip_inventory(inv)
    real(c_double), intent(out) :: inv(nl)
    call check(cgo_inventory(eng, mem(1), inv), 'cgo_inventory')
  end subroutine cgenie_hip_inventory

  subroutine cgenie_hip_diag(d)
    type(cgo_diag), intent(out) :: d
    call check(cgo_get_diag(eng, mem(1), d), 'cgo_get_diag')
  end subroutine cgenie_hip_diag

  ! tstepo on the device state (goldstein.f90:2280-2432)
  subroutine tstepo()
    call check(cgo_tstepo(eng, mem, 1_c_int32_t, 1_c_int32_t), 'cgo_tstepo')
  end subroutine tstepo

  ! the tracer part of step_goldstein: tstepo, then go_ts = ts, go_ts1 = ts1
  ! (goldstein.f90:235, 443-444; ts1 = ts after tstepo, 2423-2431)
  subroutine step_goldstein_tracers(go_ts, go_ts1)
    real(c_double), intent(inout) :: go_ts(nl, nx, ny, nz), go_ts1(nl, nx, ny, nz)
    call tstepo()
    if (mirror) call cgenie_hip_download(go_ts, go_ts1)
  end subroutine step_goldstein_tracers

  ! step_goldstein (goldstein.f90:17-478) for the tracer path.  The surface
  ! boundary slot ts(:,:,:,maxk+1) is built from the heat and freshwater
  ! fluxes as at 149-170 (no hosing or freshwater anomaly); the state comes
  ! from go_ts when mirroring is on (176-177: the host's go_ts is the state;
  ! go_ts1 equals go_ts between steps of the genie loop) and stays on the
  ! device otherwise.  Velocities are the go_u given (3,maxi,maxj,maxk; the
  ! periodic western face and the closed southern face of the engine's
  ! u(3,0:maxi,0:maxj,maxk) follow from it), handed to the engine when they
  ! change.  After tstepo: go_ts, go_ts1, go_rho and go_cost, the surface T,
  ! S and velocities, albedo and the energy and water diagnostics (427-478).
  ! With the velocity solve on (initialise_goldstein_momentum_hip) the
  ! velocities come from the wind stresses instead (102-107, 198-233) and
  ! go_u, go_tau and go_psi are handed back (444-450); go_mldta is left as
  ! given (the mixed-layer scheme, imld, is off).
  subroutine step_goldstein(istep, latent_ocn, sensible_ocn, netsolar_ocn, netlong_ocn, fx0sic_ocn, evap_ocn, &
       pptn_ocn, runoff_ocn, fwsic_ocn, stressxu_ocn, stressyu_ocn, stressxv_ocn, stressyv_ocn, tsval_ocn, &
       ssval_ocn, usval_ocn, vsval_ocn, albedo_ocn, test_energy_ocean, test_water_ocean, go_ts, go_ts1, go_cost, &
       go_u, go_tau, go_psi, go_mldta, go_rho)
    integer, intent(in) :: istep
    real(c_double), dimension(:,:), intent(inout) :: latent_ocn, sensible_ocn, netsolar_ocn, netlong_ocn, &
         fx0sic_ocn, evap_ocn, pptn_ocn, runoff_ocn, fwsic_ocn, stressxu_ocn, stressyu_ocn, stressxv_ocn, &
         stressyv_ocn, tsval_ocn, ssval_ocn, usval_ocn, vsval_ocn, albedo_ocn
    real(c_double), intent(out) :: test_energy_ocean, test_water_ocean
    real(c_double), dimension(:,:,:,:), intent(inout) :: go_ts, go_ts1
    real(c_double), dimension(:,:), intent(inout) :: go_cost, go_mldta
    real(c_double), intent(inout) :: go_u(:,:,:,:), go_tau(:,:,:), go_psi(:,:), go_rho(:,:,:)
    real(c_double) :: fx0neto(nx, ny), fwfxneto(nx, ny), tot_energy, tot_water
    real(c_double), allocatable :: u(:,:,:,:), psi(:,:)
    real(c_double) :: den
    integer :: i, j
    logical :: wet

    if (.not. c_associated(eng)) call check(CGO_E_STATE, 'step_goldstein: initialise_goldstein_hip first')
    if (istep < 0) call check(CGO_E_ARG, 'step_goldstein: negative istep')
    if (sfx_on) then
       ! surflux_hip formed the slot on the device; the sea-ice model's heat
       ! and freshwater, which it hands over after surflux, join it here
       if (any(fx0sic_ocn /= 0.0_dp) .or. any(fwsic_ocn /= 0.0_dp)) then
          sic_c(:, :, 1) = fx0sic_ocn
          sic_c(:, :, 2) = fwsic_ocn
          call check(cgo_surface_slot(eng, mem(1), c_loc(sic_c(1, 1, 1)), c_loc(sic_c(1, 1, 2))), 'cgo_surface_slot')
       end if
    else
       ! net heat and freshwater fluxes into the surface slot (149-170)
       fx0neto = netsolar_ocn + sensible_ocn + netlong_ocn + latent_ocn + fx0sic_ocn
       fwfxneto = pptn_ocn + evap_ocn + runoff_ocn + fwsic_ocn
       fwfxneto = fwfxneto * mm2m
       sflux_c = 0.0_dp
       sflux_c(1, :, :) = -fx0neto * rfluxsc
       if (nl >= 2) sflux_c(2, :, :) = fwfxneto * rpmesco()
       if (any(fx0neto /= 0.0_dp) .or. any(fwfxneto /= 0.0_dp)) then
          call check(cgo_set_sflux(eng, mem(1), c_loc(sflux_c)), 'cgo_set_sflux')
          sflux_on = .true.
       else if (sflux_on) then
          call check(cgo_set_sflux(eng, mem(1), c_null_ptr), 'cgo_set_sflux')
          sflux_on = .false.
       end if
    end if
    if (mom_on) then
       ! the wind stresses, when they differ from the engine's (102-107)
       call set_stress(stressxu_ocn, stressyu_ocn, stressxv_ocn, stressyv_ocn)
    else
       ! velocities, when the caller's differ from the engine's
       ! (the engine reads the interior faces only; the west face of i = 1 is the
       ! east face of i = maxi)
       allocate (u(3, 0:nx, 0:ny, nz))
       u = 0.0_dp
       u(:, 1:nx, 1:ny, :) = go_u(:, 1:nx, 1:ny, 1:nz)
       u(:, 0, 1:ny, :) = u(:, nx, 1:ny, :)
       if (any(u(:, 1:nx, 1:ny, :) /= u_last(:, 1:nx, 1:ny, :))) then
          call check(cgo_set_velocity(eng, mem(1), u), 'cgo_set_velocity')
          u_last = u
       end if
    end if
    if (mirror) then
       call check(cgo_upload_state(eng, mem(1), go_ts), 'cgo_upload_state')
       call check(cgo_set_cost(eng, mem(1), go_cost), 'cgo_set_cost')
    end if
    call tstepo()
    if (mom_on) then
       ! the solved velocities, tau and psi (444-450: go_psi = 1592.5 psi)
       call check(cgo_download_velocity(eng, mem(1), go_u), 'cgo_download_velocity')
       allocate (psi(0:nx, 0:ny))
       call check(cgo_download_psi(eng, mem(1), psi), 'cgo_download_psi')
       go_psi = 1592.5_dp * psi
       den = rh0sc * dsc_s * usc * fsc
       go_tau(1, :, :) = scf_s * stressxu_ocn / den / dzz_s * dzz_s
       go_tau(2, :, :) = scf_s * stressyv_ocn / den / dzz_s * dzz_s
    end if
    usval_ocn = go_u(1, 1:nx, 1:ny, nz)
    vsval_ocn = go_u(2, 1:nx, 1:ny, nz)
    if (.not. mirror) return
    call cgenie_hip_download(go_ts, go_ts1)
    call check(cgo_download_rho(eng, mem(1), go_rho), 'cgo_download_rho')
    call check(cgo_download_cost(eng, mem(1), go_cost), 'cgo_download_cost')
    do j = 1, ny
       do i = 1, nx
          wet = k1_c(i, j) <= nz
          tsval_ocn(i, j) = merge(go_ts(1, i, j, nz), 0.0_c_double, wet)
          ssval_ocn(i, j) = merge(go_ts(2, i, j, nz), 0.0_c_double, wet)
          albedo_ocn(i, j) = merge(albocn, 0.0_c_double, wet)
       end do
    end do
    call energy_water(go_ts, tot_energy, tot_water)
    test_energy_ocean = tot_energy - ini_energy
    test_water_ocean = tot_water - ini_water
  end subroutine step_goldstein

  pure function rpmesco() result(r)
    real(c_double) :: r
    r = rsc * saln0 / (dsc_p * usc)   ! goldstein.f90:904 (computed once per call: the reference's is a variable)
  end function rpmesco

  ! step_biogem (biogem.f90:528-547).  dum_dts is the BIOGEM step in seconds
  ! (loc_dtyr = dum_dts / conv_yr_s, 622-624); it must be the step
  ! initialise_biogem_hip was given.  With mirroring on, the interface arrays
  ! BIOGEM writes are filled as at 1719-1754: dum_sfcocn1(io,i,j) the
  ! bottom-water composition after this step's biology (ocn + the step's
  ! anomaly at k1, Ca, B, F and SO4 from salinity), dum_sfxsed1(is,i,j) the
  ! rain onto the sea floor in mol m-2 s-1 (the recalcitrant fractions
  ! scaled by dtyr / dts).  Air-sea exchange and sediment / weathering inputs
  ! are not part of this path: dum_sfxatm1, dum_sfxocn1 and dum_sfxsumrok1
  ! are left as given.
  subroutine step_biogem(dum_dts, dum_genie_clock, dum_sfcatm1, dum_sfxatm1, dum_sfcocn1, dum_sfxocn1, &
       dum_sfcsed1, dum_sfxsed1, dum_sfxsumrok1)
    real(c_double), intent(in) :: dum_dts
    integer(c_int64_t), intent(in) :: dum_genie_clock
    real(c_double), intent(in), dimension(:,:,:) :: dum_sfcatm1, dum_sfcsed1
    real(c_double), intent(inout), dimension(:,:,:) :: dum_sfxatm1, dum_sfxocn1, dum_sfxsed1, dum_sfxsumrok1
    real(c_double), intent(out), dimension(:,:,:) :: dum_sfcocn1
    real(c_double) :: dtyr, S, A
    real(c_double), allocatable :: ocn(:,:,:,:), vd(:,:,:,:), settle(:,:,:), fatm(:,:,:)
    integer :: i, j, kb, q, l, io, is

    if (.not. (dum_dts > 0.0_c_double)) call check(CGO_E_ARG, 'step_biogem: dum_dts must be positive')
    dtyr = dum_dts / conv_yr_s
    if (abs(dtyr - bgc_dtyr) > 1.0E-12_dp * bgc_dtyr) &
         call check(CGO_E_ARG, 'step_biogem: dum_dts differs from the step initialise_biogem_hip was given')
    call check(cgo_bio_step(eng, mem, 1_c_int32_t), 'cgo_bio_step')
    dum_sfcocn1 = 0.0_dp
    if (.not. mirror) return
    allocate (ocn(nl, nx, ny, nz), vd(nl, nx, ny, nz), settle(5, nx, ny), fatm(2, nx, ny))
    call check(cgo_bgc_download(eng, mem(1), 0_c_int32_t, ocn), 'cgo_bgc_download ocn')
    if (airsea) call check(cgo_bgc_download(eng, mem(1), 7_c_int32_t, fatm), 'cgo_bgc_download air-sea flux')
    call check(cgo_bgc_download(eng, mem(1), 3_c_int32_t, vd), 'cgo_bgc_download anomalies')
    call check(cgo_bgc_download(eng, mem(1), 5_c_int32_t, settle), 'cgo_bgc_download rain')
    do j = 1, ny
       do i = 1, nx
          kb = k1_c(i, j)
          if (kb > nz) cycle
          ! (2) bottom-water tracers: T, S and the carried species
          do l = 1, 2
             if (l <= size(dum_sfcocn1, 1)) dum_sfcocn1(l, i, j) = ocn(l, i, j, kb) + vd(l, i, j, kb)
          end do
          do q = 1, 6
             if (bgc_slots(q) <= 0) cycle
             io = io_of_slot(q)
             if (io <= size(dum_sfcocn1, 1)) dum_sfcocn1(io, i, j) = ocn(bgc_slots(q), i, j, kb) + &
                  vd(bgc_slots(q), i, j, kb)
          end do
          if (size(dum_sfcocn1, 1) >= io_SO4 .and. size(dum_sfcocn1, 1) >= 2) then
             S = dum_sfcocn1(2, i, j)
             dum_sfcocn1(io_Ca, i, j) = max(0.01028_dp * S / 35.0_dp, nullsmall)    ! gem_carbchem.f90:834-890
             dum_sfcocn1(io_B, i, j) = max(0.000416_dp * S / 35.0_dp, nullsmall)
             dum_sfcocn1(io_SO4, i, j) = max(0.02824_dp * S / 35.0_dp, nullsmall)
             dum_sfcocn1(io_F, i, j) = max(0.00007_dp * S / 35.0_dp, nullsmall)
          end if
          ! (3) ocean -> sediment flux (mol m-2 s-1); the bottom cell's area
          ! as sub_init_phys_ocn forms it (biogem_data.f90:1120-1128)
          A = 2.0_dp * const_pi * (const_rEarth * const_rEarth) * (1.0_dp / nx) * (jtab(j, 9) - jtab(j - 1, 9))
          ! (1) ocean -> atmosphere flux of pCO2 and pO2 in mol m-2 s-1
          ! (biogem.f90:1721-1727); ATCHEM itself runs on the device
          if (airsea .and. size(dum_sfxatm1, 1) >= ia_pO2) then
             dum_sfxatm1(ia_pCO2, i, j) = (1.0_dp / A) * conv_s_yr * fatm(1, i, j)
             dum_sfxatm1(ia_pO2, i, j) = (1.0_dp / A) * conv_s_yr * fatm(2, i, j)
          end if
          do q = 1, 5
             is = is_of(q)
             if (is > size(dum_sfxsed1, 1)) cycle
             if (q >= 4) then
                dum_sfxsed1(is, i, j) = settle(q, i, j) * (1.0_dp / dum_dts) * dtyr
             else
                dum_sfxsed1(is, i, j) = (1.0_dp / A) * settle(q, i, j) * (1.0_dp / dum_dts)
             end if
          end do
       end do
    end do
  end subroutine step_biogem

  ! biogem_tracercoupling(dum_ts, dum_ts1) (biogem.f90:1885-1890)
  subroutine biogem_tracercoupling(dum_ts, dum_ts1)
    real(c_double), intent(inout) :: dum_ts(nl, nx, ny, nz), dum_ts1(nl, nx, ny, nz)
    call check(cgo_tracercoupling(eng, mem, 1_c_int32_t), 'cgo_tracercoupling')
    if (mirror) call cgenie_hip_download(dum_ts, dum_ts1)
  end subroutine biogem_tracercoupling

  subroutine end_cgenie_hip()
    integer(c_int) :: rc
    if (.not. c_associated(eng)) return
    rc = cgo_destroy(eng)
    eng = c_null_ptr
    if (allocated(k1_c)) deallocate (k1_c, jtab, ktab)
    if (allocated(u_last)) deallocate (u_last, sflux_c)
    if (allocated(sfx_in)) deallocate (sfx_in, sfx_out, sic_c)
    sfx_on = .false.
  end subroutine end_cgenie_hip

end module cgenie_hip
