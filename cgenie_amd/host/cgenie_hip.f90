! cgenie_hip.f90 -- drop-in Fortran host module for the tracer path of one
! cGENIE ocean step on the MI355X engine.
!
! It keeps the reference's calling pattern: module state set up once (as
! initialise_goldstein / initialise_biogem do, goldstein.f90:906-1075,
! biogem_data.f90:1098-1140), then per ocean step the calls that genie.f90
! makes (genie.f90:306-313 GOLDSTEIN, 360-381 BIOGEM):
!
!   tstepo()                              goldstein.f90:2280-2432, called at 235
!   step_goldstein_tracers(go_ts, go_ts1) the tracer part of step_goldstein
!                                         (goldstein.f90:17-38, 167-187, 235, 443-444)
!   step_biogem(dum_dts)                  biogem.f90:528-547 (biology, remineralisation
!                                         and surface chemistry of the carried species)
!   biogem_tracercoupling(dum_ts, dum_ts1) biogem.f90:1885-1890
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
  public :: cgenie_hip_set_sflux

  type(c_ptr), save :: eng = c_null_ptr
  integer(c_int32_t), save :: grid_id = -1, mem(1) = -1
  integer, save :: nx = 0, ny = 0, nz = 0, nl = 0
  logical, save :: mirror = .true.
  ! arrays whose addresses the grid descriptor holds during cgo_grid_create
  integer(c_int32_t), allocatable, target, save :: k1_c(:,:)
  real(c_double), allocatable, target, save :: jtab(:,:), ktab(:,:)

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
  end subroutine initialise_goldstein_hip

  ! BIOGEM state of the member from its current go_ts.  slots(1:6) are the
  ! 1-based go_ts positions of DIC, ALK, PO4, O2, DOM_C, DOM_P (0 = absent);
  ! light = solfor/solar_constant and mld (m) are (maxi,maxj).
  subroutine initialise_biogem_hip(slots, dtyr, light, mld, carb_tol)
    integer, intent(in) :: slots(6)
    real(c_double), intent(in) :: dtyr
    real(c_double), intent(in) :: light(nx, ny), mld(nx, ny)
    real(c_double), intent(in), optional :: carb_tol
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
    call check(cgo_bgc_init(eng, mem(1), bp, light, mld), 'cgo_bgc_init')
  end subroutine initialise_biogem_hip

  ! the surface boundary slot ts(:,1:maxi,1:maxj,maxk+1) (goldstein.f90:167-170)
  subroutine cgenie_hip_set_sflux(sflux)
    real(c_double), intent(in), target :: sflux(nl, nx, ny)
    call check(cgo_set_sflux(eng, mem(1), c_loc(sflux)), 'cgo_set_sflux')
  end subroutine cgenie_hip_set_sflux

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

  ! step_biogem (biogem.f90:528-547); dum_dts is the BIOGEM step of the
  ! parameters given to initialise_biogem_hip
  subroutine step_biogem(dum_dts)
    real(c_double), intent(in) :: dum_dts
    if (.not. (dum_dts > 0.0_c_double)) call check(CGO_E_ARG, 'step_biogem: dum_dts must be positive')
    call check(cgo_bio_step(eng, mem, 1_c_int32_t), 'cgo_bio_step')
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
  end subroutine end_cgenie_hip

end module cgenie_hip
