! ref_gold_shim.f90 -- a C entry point over the reference's own path integral
! round an island (goldstein_lib.f90:186-241, subroutine island), compiled in
! place from the reference source by oracle/Makefile (target "ref").  TEST
! INFRASTRUCTURE ONLY: used by tests/ to pin the C oracle's island integral
! to the reference arithmetic.
!
! island() reads GOLDSTEIN's module arrays; the shim allocates them with the
! bounds goldstein.f90 gives them (ku 657, mk 659, dz 682, s 688, c 690,
! tau 694, drag 696, sv 708, rh 722, rcv 766, dsv 678, bp 778, sbp 780, the
! island paths 1522-1528), copies the caller's values in and calls island().
! goldstein.f90 itself needs the netCDF module and is not built.
module ref_gold_shim
  use iso_c_binding
  use goldstein_lib
  implicit none
contains

  subroutine ref_gold_island(mi, mj, mkk, nisl, mpii, npi_in, lp_in, ip_in, jp_in, &
       & s_in, sv_in, c_in, rcv_in, dsv_in, dz_in, dphi_in, drag_in, rh_in, tau_in, &
       & bp_in, sbp_in, ku_in, mk_in, ub_in, isl, indj, out) bind(C, name="ref_gold_island")
    integer(c_int), value :: mi, mj, mkk, nisl, mpii, isl, indj
    integer(c_int), intent(in) :: npi_in(nisl), lp_in(mpii, nisl), ip_in(mpii, nisl), jp_in(mpii, nisl)
    real(c_double), intent(in) :: s_in(0:mj), sv_in(0:mj), c_in(0:mj), rcv_in(1:mj-1), dsv_in(1:mj-1)
    real(c_double), intent(in) :: dz_in(mkk)
    real(c_double), value :: dphi_in
    real(c_double), intent(in) :: drag_in(2, mi+1, mj), rh_in(3, 0:mi+1, 0:mj+1), tau_in(2, mi, mj)
    real(c_double), intent(in) :: bp_in(mi+1, mj, mkk), sbp_in(mi+1, mj)
    integer(c_int), intent(in) :: ku_in(2, mi, mj), mk_in(mi+1, mj)
    real(c_double), intent(in) :: ub_in(2, 0:mi+1, 0:mj)
    real(c_double), intent(out) :: out
    real :: e

    maxi = mi; maxj = mj; maxk = mkk; isles = nisl; mpi = mpii
    if (allocated(s)) deallocate (s, sv, c, rcv, dsv, dz, drag, rh, tau, bp, sbp, ku, mk, npi, lpisl, ipisl, jpisl)
    allocate (s(0:mj), sv(0:mj), c(0:mj), rcv(1:mj-1), dsv(1:mj-1), dz(mkk))
    allocate (drag(2, mi+1, mj), rh(3, 0:mi+1, 0:mj+1), tau(2, mi, mj), bp(mi+1, mj, mkk), sbp(mi+1, mj))
    allocate (ku(2, mi, mj), mk(mi+1, mj), npi(nisl), lpisl(mpii, nisl), ipisl(mpii, nisl), jpisl(mpii, nisl))
    s = s_in; sv = sv_in; c = c_in; rcv = rcv_in; dsv = dsv_in; dz = dz_in; dphi = dphi_in
    drag = drag_in; rh = rh_in; tau = tau_in; bp = bp_in; sbp = sbp_in; ku = ku_in; mk = mk_in
    npi = npi_in; lpisl = lp_in; ipisl = ip_in; jpisl = jp_in
    call island(ub_in, e, isl, indj)
    out = e
  end subroutine ref_gold_island

end module ref_gold_shim
