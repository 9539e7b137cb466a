! cgo_api.f90 -- Fortran interface of libcgocn.so (the C ABI of
! include/cgocn.h), written with iso_c_binding so that Fortran code of the
! cGENIE kind (REAL*8 arrays in the reference's own layouts) calls the MI355X
! engine directly.  Array arguments are passed by address in Fortran
! (column-major) order; the engine reads and writes them only inside the call.
module cgo_api
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: CGO_OK = 0, CGO_E_ARG = -1, CGO_E_HIP = -2, CGO_E_STATE = -3, CGO_E_NOMEM = -4
  integer(c_int32_t), parameter :: CGO_FP64 = 0, CGO_FP32C = 1

  ! grid metrics of initialise_goldstein (goldstein.f90:906-1075): per-j
  ! arrays (0:maxj+1), per-k arrays (0:maxk+1), k1(0:maxi+1,0:maxj+1)
  type, bind(C) :: cgo_grid_desc
     integer(c_int32_t) :: maxi, maxj, maxk
     type(c_ptr) :: k1
     real(c_double) :: dphi, rdphi
     type(c_ptr) :: rc, rc2, cv, cv2, rds, dsv, rdsv, ds, sv
     type(c_ptr) :: dt, rdz, rdza, dza, ssmax, dz
     real(c_double) :: ec(5)
     real(c_double) :: dsc
  end type cgo_grid_desc

  type, bind(C) :: cgo_member_params
     real(c_double) :: diff(2)
     integer(c_int32_t) :: diso, carry
  end type cgo_member_params

  ! BIOGEM parameters; slots are 0-based tracer positions, -1 = absent
  type, bind(C) :: cgo_bgc_params
     integer(c_int32_t) :: slot_DIC, slot_ALK, slot_PO4, slot_O2, slot_DOM_C, slot_DOM_P
     integer(c_int32_t) :: caco3
     real(c_double) :: dtyr, k0_PO4, c0_PO4, DOMfrac, DOMlifetime
     real(c_double) :: red_POP_POC, red_POP_PO2, red_PON_ALK, red_POP_PON, red_POC_CaCO3, red_POC_CaCO3_pP
     real(c_double) :: POC_frac2, POC_dfrac2, POC_c0frac2, POC_eL1, POC_eL2
     real(c_double) :: CaCO3_frac2, CaCO3_eL1, CaCO3_eL2
     real(c_double) :: sinkrate, remin_k_O2, remin_c0_O2, carb_tol, dsc
     integer(c_int32_t) :: carb_maxit
     ! air-sea exchange of CO2 and O2 with the well-mixed ATCHEM atmosphere
     integer(c_int32_t) :: airsea
     real(c_double) :: gastransfer_a, atm_pCO2, atm_pO2
  end type cgo_bgc_params

  ! GOLDSTEIN's velocity solve of a grid: the per-j metrics s, c, rcv, rds2
  ! (0:maxj+1), the .psiles labels (maxi,0:maxj), island count, path lengths
  ! and (lpisl, ipisl, jpisl) triples, drag in days and relaxation
  type, bind(C) :: cgo_momentum_desc
     type(c_ptr) :: s, c, rcv, rds2
     type(c_ptr) :: labels
     integer(c_int32_t) :: isles
     type(c_ptr) :: npi, paths
     real(c_double) :: adrag, rel
  end type cgo_momentum_desc

  ! EMBM's surface fluxes of a grid (cgo_grid_set_surflux): initialise_embm's
  ! tables, each c_loc of the reference's array, and its namelist constants
  type, bind(C) :: cgo_surflux_desc
     type(c_ptr) :: solfor            ! solfor(maxj, nyear)
     integer(c_int32_t) :: nyear
     type(c_ptr) :: albcl, ca, pmeadj ! (maxi, maxj); pmeadj may be c_null_ptr
     type(c_ptr) :: iroff, jroff      ! (maxi, maxj) integer(c_int32_t)
     integer(c_int32_t) :: igrid, polar_avg, radfor_atchem
     real(c_double) :: scf, rmax, delf2x, rdtdim, rsictscsf, olr_adj, olr_adj0, t_eqm
     real(c_double) :: albsic_min, albsic_max, sich_max, rate_co2, rate_ch4, rate_n2o
     real(c_double) :: co2, ch4, n2o
  end type cgo_surflux_desc

  ! the surface fluxes' outputs (include/cgocn.h CGO_SF_*, 0-based planes)
  integer, parameter :: CGO_SF_N = 28
  integer, parameter :: CGO_SF_FXLHO = 0, CGO_SF_FXSHO = 1, CGO_SF_FXSWO = 2, CGO_SF_FXLWO = 3, CGO_SF_EVAP_OCN = 4, &
       CGO_SF_PPTN_OCN = 5, CGO_SF_RUNOFF_OCN = 6, CGO_SF_RUNOFF_LAND = 7, CGO_SF_FXLHA = 8, CGO_SF_FXSHA = 9, &
       CGO_SF_FXSWA = 10, CGO_SF_FXLWA = 11, CGO_SF_EVAP_ATM = 12, CGO_SF_PPTN_ATM = 13, CGO_SF_DTHSIC = 14, &
       CGO_SF_DTAREASIC = 15, CGO_SF_ALBEDO = 16, CGO_SF_ALBICE = 17

  type, bind(C) :: cgo_diag
     integer(c_int64_t) :: steps
     integer(c_int32_t) :: limps, carb_fail
     integer(c_int64_t) :: carb_iters
     real(c_double) :: dmax, cost, smean
  end type cgo_diag

  interface
     integer(c_int) function cgo_create(device, out) bind(C, name="cgo_create")
       import :: c_int, c_int32_t, c_ptr
       integer(c_int32_t), value :: device
       type(c_ptr), intent(out) :: out
     end function cgo_create

     integer(c_int) function cgo_destroy(e) bind(C, name="cgo_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: e
     end function cgo_destroy

     integer(c_int) function cgo_last_error(e, buf, n) bind(C, name="cgo_last_error")
       import :: c_int, c_ptr, c_char, c_size_t
       type(c_ptr), value :: e
       character(kind=c_char), intent(out) :: buf(*)
       integer(c_size_t), value :: n
     end function cgo_last_error

     integer(c_int) function cgo_sync(e) bind(C, name="cgo_sync")
       import :: c_int, c_ptr
       type(c_ptr), value :: e
     end function cgo_sync

     integer(c_int) function cgo_grid_create(e, g, grid_id) bind(C, name="cgo_grid_create")
       import :: c_int, c_ptr, c_int32_t, cgo_grid_desc
       type(c_ptr), value :: e
       type(cgo_grid_desc), intent(in) :: g
       integer(c_int32_t), intent(out) :: grid_id
     end function cgo_grid_create

     integer(c_int) function cgo_member_create(e, grid_id, maxl, precision, p, member_id) &
          bind(C, name="cgo_member_create")
       import :: c_int, c_ptr, c_int32_t, cgo_member_params
       type(c_ptr), value :: e
       integer(c_int32_t), value :: grid_id, maxl, precision
       type(cgo_member_params), intent(in) :: p
       integer(c_int32_t), intent(out) :: member_id
     end function cgo_member_create

     integer(c_int) function cgo_member_destroy(e, m) bind(C, name="cgo_member_destroy")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: e
       integer(c_int32_t), value :: m
     end function cgo_member_destroy

     integer(c_int64_t) function cgo_member_bytes(e, grid_id, maxl, precision, bgc) bind(C, name="cgo_member_bytes")
       import :: c_int64_t, c_ptr, c_int32_t
       type(c_ptr), value :: e
       integer(c_int32_t), value :: grid_id, maxl, precision, bgc
     end function cgo_member_bytes

     integer(c_int) function cgo_set_velocity(e, m, u) bind(C, name="cgo_set_velocity")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: e
       integer(c_int32_t), value :: m
       real(c_double), intent(in) :: u(*)
     end function cgo_set_velocity

     integer(c_int) function cgo_grid_set_momentum(e, grid_id, d) bind(C, name="cgo_grid_set_momentum")
       import :: c_int, c_ptr, c_int32_t, cgo_momentum_desc
       type(c_ptr), value :: e
       integer(c_int32_t), value :: grid_id
       type(cgo_momentum_desc), intent(in) :: d
     end function cgo_grid_set_momentum

     ! the four wind-stress arrays of step_goldstein, (maxi,maxj) each
     integer(c_int) function cgo_set_wind(e, m, sxu, syu, sxv, syv, scf) bind(C, name="cgo_set_wind")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: e
       integer(c_int32_t), value :: m
       real(c_double), intent(in) :: sxu(*), syu(*), sxv(*), syv(*)
       real(c_double), value :: scf
     end function cgo_set_wind

     ! go_u(3,maxi,maxj,maxk)
     integer(c_int) function cgo_download_velocity(e, m, go_u) bind(C, name="cgo_download_velocity")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: e
       integer(c_int32_t), value :: m
       real(c_double), intent(out) :: go_u(*)
     end function cgo_download_velocity

     ! psi(0:maxi,0:maxj), non-dimensional
     integer(c_int) function cgo_download_psi(e, m, psi) bind(C, name="cgo_download_psi")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: e
       integer(c_int32_t), value :: m
       real(c_double), intent(out) :: psi(*)
     end function cgo_download_psi

     integer(c_int) function cgo_grid_set_surflux(e, grid_id, d) bind(C, name="cgo_grid_set_surflux")
       import :: c_int, c_ptr, c_int32_t, cgo_surflux_desc
       type(c_ptr), value :: e
       integer(c_int32_t), value :: grid_id
       type(cgo_surflux_desc), intent(in) :: d
     end function cgo_grid_set_surflux

     ! the member's air and sea-ice state, each c_loc of a (maxi,maxj) array
     ! or c_null_ptr (zero); atemp c_null_ptr switches the fluxes off
     integer(c_int) function cgo_set_surface(e, m, atemp, ashum, sich, sica, tice, fx0sic, fwsic) &
          bind(C, name="cgo_set_surface")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: e
       integer(c_int32_t), value :: m
       type(c_ptr), value :: atemp, ashum, sich, sica, tice, fx0sic, fwsic
     end function cgo_set_surface

     integer(c_int) function cgo_surflux(e, members, n) bind(C, name="cgo_surflux")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: e
       integer(c_int32_t), intent(in) :: members(*)
       integer(c_int32_t), value :: n
     end function cgo_surflux

     ! the slot again with new sea-ice fluxes, each c_loc of (maxi,maxj) or c_null_ptr
     integer(c_int) function cgo_surface_slot(e, m, fx0sic, fwsic) bind(C, name="cgo_surface_slot")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: e
       integer(c_int32_t), value :: m
       type(c_ptr), value :: fx0sic, fwsic
     end function cgo_surface_slot

     ! out(maxi,maxj,CGO_SF_N); ashum, tice (maxi,maxj); gases(4)
     integer(c_int) function cgo_download_surflux(e, m, out, ashum, tice, gases) bind(C, name="cgo_download_surflux")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: e
       integer(c_int32_t), value :: m
       type(c_ptr), value :: out, ashum, tice, gases
     end function cgo_download_surflux

     ! sflux: c_loc of ts(:,1:maxi,1:maxj,maxk+1) values, or c_null_ptr
     integer(c_int) function cgo_set_sflux(e, m, sflux) bind(C, name="cgo_set_sflux")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: e
       integer(c_int32_t), value :: m
       type(c_ptr), value :: sflux
     end function cgo_set_sflux

     integer(c_int) function cgo_upload_state(e, m, go_ts) bind(C, name="cgo_upload_state")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: e
       integer(c_int32_t), value :: m
       real(c_double), intent(in) :: go_ts(*)
     end function cgo_upload_state

     integer(c_int) function cgo_download_state(e, m, go_ts) bind(C, name="cgo_download_state")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: e
       integer(c_int32_t), value :: m
       real(c_double), intent(inout) :: go_ts(*)
     end function cgo_download_state

     integer(c_int) function cgo_download_rho(e, m, rho) bind(C, name="cgo_download_rho")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: e
       integer(c_int32_t), value :: m
       real(c_double), intent(out) :: rho(*)
     end function cgo_download_rho

     integer(c_int) function cgo_inventory(e, m, inv) bind(C, name="cgo_inventory")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: e
       integer(c_int32_t), value :: m
       real(c_double), intent(out) :: inv(*)
     end function cgo_inventory

     integer(c_int) function cgo_get_diag(e, m, d) bind(C, name="cgo_get_diag")
       import :: c_int, c_ptr, c_int32_t, cgo_diag
       type(c_ptr), value :: e
       integer(c_int32_t), value :: m
       type(cgo_diag), intent(out) :: d
     end function cgo_get_diag

     integer(c_int) function cgo_set_cost(e, m, cost) bind(C, name="cgo_set_cost")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: e
       integer(c_int32_t), value :: m
       real(c_double), intent(in) :: cost(*)
     end function cgo_set_cost

     integer(c_int) function cgo_download_cost(e, m, cost) bind(C, name="cgo_download_cost")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: e
       integer(c_int32_t), value :: m
       real(c_double), intent(out) :: cost(*)
     end function cgo_download_cost

     integer(c_int) function cgo_inventory_noalloc(e, m, inv) bind(C, name="cgo_inventory_noalloc")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: e
       integer(c_int32_t), value :: m
       real(c_double), intent(out) :: inv(*)
     end function cgo_inventory_noalloc

     subroutine cgo_bgc_default_params(p, dtyr) bind(C, name="cgo_bgc_default_params")
       import :: c_double, cgo_bgc_params
       type(cgo_bgc_params), intent(out) :: p
       real(c_double), value :: dtyr
     end subroutine cgo_bgc_default_params

     integer(c_int) function cgo_bgc_init(e, m, p, light, mld) bind(C, name="cgo_bgc_init")
       import :: c_int, c_ptr, c_int32_t, c_double, cgo_bgc_params
       type(c_ptr), value :: e
       integer(c_int32_t), value :: m
       type(cgo_bgc_params), intent(in) :: p
       real(c_double), intent(in) :: light(*), mld(*)
     end function cgo_bgc_init

     integer(c_int) function cgo_bgc_download(e, m, what, out) bind(C, name="cgo_bgc_download")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: e
       integer(c_int32_t), value :: m, what
       real(c_double), intent(out) :: out(*)
     end function cgo_bgc_download

     integer(c_int) function cgo_set_windspeed(e, m, wspeed) bind(C, name="cgo_set_windspeed")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: e
       integer(c_int32_t), value :: m
       real(c_double), intent(in) :: wspeed(*)
     end function cgo_set_windspeed

     integer(c_int) function cgo_step(e, members, n, nsteps) bind(C, name="cgo_step")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: e
       integer(c_int32_t), intent(in) :: members(*)
       integer(c_int32_t), value :: n, nsteps
     end function cgo_step

     integer(c_int) function cgo_tstepo(e, members, n, nsteps) bind(C, name="cgo_tstepo")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: e
       integer(c_int32_t), intent(in) :: members(*)
       integer(c_int32_t), value :: n, nsteps
     end function cgo_tstepo

     integer(c_int) function cgo_bio_step(e, members, n) bind(C, name="cgo_bio_step")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: e
       integer(c_int32_t), intent(in) :: members(*)
       integer(c_int32_t), value :: n
     end function cgo_bio_step

     integer(c_int) function cgo_tracercoupling(e, members, n) bind(C, name="cgo_tracercoupling")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: e
       integer(c_int32_t), intent(in) :: members(*)
       integer(c_int32_t), value :: n
     end function cgo_tracercoupling

     integer(c_int) function cgo_graph_capture(e, members, n, nsteps, graph_id) bind(C, name="cgo_graph_capture")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: e
       integer(c_int32_t), intent(in) :: members(*)
       integer(c_int32_t), value :: n, nsteps
       integer(c_int32_t), intent(out) :: graph_id
     end function cgo_graph_capture

     integer(c_int) function cgo_graph_launch(e, graph_id, nreps) bind(C, name="cgo_graph_launch")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: e
       integer(c_int32_t), value :: graph_id, nreps
     end function cgo_graph_launch

     integer(c_int) function cgo_graph_destroy(e, graph_id) bind(C, name="cgo_graph_destroy")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: e
       integer(c_int32_t), value :: graph_id
     end function cgo_graph_destroy

     integer(c_int) function cgo_set_tile(e, precision, tracer_block) bind(C, name="cgo_set_tile")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: e
       integer(c_int32_t), value :: precision, tracer_block
     end function cgo_set_tile

     integer(c_int) function cgo_set_flux_cols(e, precision, cols) bind(C, name="cgo_set_flux_cols")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: e
       integer(c_int32_t), value :: precision, cols
     end function cgo_set_flux_cols

     integer(c_int) function cgo_set_flux_ring(e, precision, planes) bind(C, name="cgo_set_flux_ring")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: e
       integer(c_int32_t), value :: precision, planes
     end function cgo_set_flux_ring

     integer(c_int) function cgo_set_flux_iso(e, precision, prepass) bind(C, name="cgo_set_flux_iso")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: e
       integer(c_int32_t), value :: precision, prepass
     end function cgo_set_flux_iso

     integer(c_int) function cgo_time_kernel(e, which, members, n, reps, avg_ms) bind(C, name="cgo_time_kernel")
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: e
       integer(c_int32_t), value :: which
       integer(c_int32_t), intent(in) :: members(*)
       integer(c_int32_t), value :: n, reps
       real(c_double), intent(out) :: avg_ms
     end function cgo_time_kernel
  end interface
end module cgo_api
