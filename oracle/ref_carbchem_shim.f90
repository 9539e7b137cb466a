! ref_carbchem_shim.f90 -- C entry points over the reference's own
! gem_carbchem module, compiled in place from the reference sources by
! oracle/Makefile (target "ref").  TEST INFRASTRUCTURE ONLY: used by tests/
! to pin the C oracle's carbonate chemistry to the reference arithmetic.
!
! Wraps: sub_calc_carbconst (gem_carbchem.f90:59-268), sub_calc_carb
! (326-546), sub_calc_carb_RF0 (553-669) and the salinity totals (834-890).
! Module controls (gem_cmn.f90:43-49, 488-489) are set from the caller.
module ref_carbchem_shim
  use iso_c_binding
  use gem_cmn
  use gem_carbchem
  implicit none
contains

  subroutine ref_set_params(tol, maxit) bind(C, name="ref_set_params")
    real(c_double), value :: tol
    integer(c_int), value :: maxit
    par_carbconstset_name = 'Mehrbach'
    par_carbchem_pH_tolerance = tol
    par_carbchem_pH_iterationmax = maxit
    ctrl_carbchem_fail = .true.
  end subroutine ref_set_params

  subroutine ref_calc_carbconst(d, t, s, cc) bind(C, name="ref_calc_carbconst")
    real(c_double), value :: d, t, s
    real(c_double), intent(inout) :: cc(n_carbconst)
    call sub_calc_carbconst(d, t, s, cc)
  end subroutine ref_calc_carbconst

  subroutine ref_salt_totals(s, out) bind(C, name="ref_salt_totals")
    real(c_double), value :: s
    real(c_double), intent(out) :: out(4)
    out(1) = fun_calc_Ca(s)
    out(2) = fun_calc_Btot(s)
    out(3) = fun_calc_SO4tot(s)
    out(4) = fun_calc_Ftot(s)
  end subroutine ref_salt_totals

  ! returns 0 on success, -1 when the reference flagged a failure
  integer(c_int) function ref_calc_carb(dic, alk, ca, po4, sio2, b, so4, f, h2s, nh4, cc, carb, carbalk) &
       & bind(C, name="ref_calc_carb")
    real(c_double), value :: dic, alk, ca, po4, sio2, b, so4, f, h2s, nh4
    real(c_double), intent(in) :: cc(n_carbconst)
    real(c_double), intent(inout) :: carb(n_carb), carbalk(n_carbalk)
    error_stop = .false.
    error_carbchem = .false.
    call sub_calc_carb(dic, alk, ca, po4, sio2, b, so4, f, h2s, nh4, cc, carb, carbalk)
    ref_calc_carb = 0
    if (error_carbchem) ref_calc_carb = -1
  end function ref_calc_carb

  subroutine ref_calc_carb_rf0(dic, alk, po4, sio2, b, so4, f, h2s, nh4, cc, carb) &
       & bind(C, name="ref_calc_carb_rf0")
    real(c_double), value :: dic, alk, po4, sio2, b, so4, f, h2s, nh4
    real(c_double), intent(in) :: cc(n_carbconst)
    real(c_double), intent(inout) :: carb(n_carb)
    call sub_calc_carb_RF0(dic, alk, po4, sio2, b, so4, f, h2s, nh4, cc, carb)
  end subroutine ref_calc_carb_rf0

end module ref_carbchem_shim
