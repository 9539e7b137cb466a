! ref_airsea_shim.f90 -- C entry points over the reference's own air-sea
! gas exchange (biogem_box.f90: sub_calc_solconst 27-40, sub_calc_pv 81-117,
! fun_calc_ocnatm_flux 123-299) and seawater density (gem_util.f90:670-688),
! compiled in place from the reference sources by oracle/Makefile (target
! "ref").  TEST INFRASTRUCTURE ONLY: used by tests/ to pin the C oracle's
! air-sea exchange to the reference arithmetic.
!
! The reference routines work on BIOGEM's module arrays; the shim sets up a
! one-cell ocean (n_i = n_j = n_k = 1) with CO2 and O2 as the selected
! atmospheric tracers, the reference's Schmidt-number and Bunsen tables
! (gem_data.f90:68-141), and calls them as step_biogem does
! (biogem.f90:1044-1046, 1605).
module ref_airsea_shim
  use iso_c_binding
  use gem_cmn
  use gem_util
  use gem_carbchem
  use gem_data
  use biogem_lib
  use biogem_box
  implicit none
  logical, save :: ready = .false.
contains

  subroutine setup()
    if (ready) return
    n_i = 1; n_j = 1; n_k = 1
    allocate (ocn(n_ocn, 1, 1, 1), carb(n_carb, 1, 1, 1), phys_ocn(n_phys_ocn, 1, 1, 1))
    allocate (phys_ocnatm(n_phys_ocnatm, 1, 1), ocnatm_airsea_pv(n_atm, 1, 1), ocnatm_airsea_solconst(n_atm, 1, 1))
    ocn = 0.0; carb = 0.0; phys_ocn = 0.0; phys_ocnatm = 0.0
    ocnatm_airsea_pv = 0.0; ocnatm_airsea_solconst = 0.0
    atm_select(:) = .false.
    atm_select(ia_T) = .true.; atm_select(ia_q) = .true.
    atm_select(ia_pCO2) = .true.; atm_select(ia_pO2) = .true.
    n_l_atm = 4
    allocate (conv_iselected_ia(n_l_atm))
    conv_iselected_ia = (/ ia_T, ia_q, ia_pCO2, ia_pO2 /)
    atm_type(:) = 0
    atm_type(ia_pCO2) = 1      ! tracer_define.atm: gases of type 1
    atm_type(ia_pO2) = 1
    ocnatm_airsea_eqm(:) = .false.
    conv_atm_ocn_i(:,:) = 0
    conv_atm_ocn_i(0, ia_pCO2) = 1; conv_atm_ocn_i(1, ia_pCO2) = io_DIC
    conv_atm_ocn_i(0, ia_pO2) = 1;  conv_atm_ocn_i(1, ia_pO2) = io_O2
    call sub_def_schmidtnumber()
    call sub_def_bunsencoefficient()
    ready = .true.
  end subroutine setup

  real(c_double) function ref_calc_rho(t, s) bind(C, name="ref_calc_rho")
    real(c_double), value :: t, s
    ref_calc_rho = fun_calc_rho(t, s)
  end function ref_calc_rho

  ! one surface cell: T (K), S, rho, layer thickness dD (m), area A (m2),
  ! wind speed (m/s), DIC, O2, carb CO2 and RF0, atmospheric pCO2 and pO2
  ! (atm), the BIOGEM step dtyr (yr) and the gas-transfer coefficient a.
  ! out(1:2) solubility of CO2, O2; out(3:4) piston velocity; out(5:6) the
  ! ocean -> atmosphere flux of CO2 and O2 (mol yr-1)
  subroutine ref_airsea_cell(t, s, rho, dd, a, wspeed, dic, o2, co2, rf0, pco2, po2, dtyr, gas_a, out) &
       & bind(C, name="ref_airsea_cell")
    real(c_double), value :: t, s, rho, dd, a, wspeed, dic, o2, co2, rf0, pco2, po2, dtyr, gas_a
    real(c_double), intent(out) :: out(6)
    real :: atm(n_atm), f(n_atm)
    call setup()
    par_gastransfer_a = gas_a
    ocn(io_T, 1, 1, 1) = t
    ocn(io_S, 1, 1, 1) = s
    ocn(io_DIC, 1, 1, 1) = dic
    ocn(io_O2, 1, 1, 1) = o2
    carb(ic_conc_CO2, 1, 1, 1) = co2
    carb(ic_RF0, 1, 1, 1) = rf0
    phys_ocn(ipo_rho, 1, 1, 1) = rho
    phys_ocn(ipo_dD, 1, 1, 1) = dd
    phys_ocnatm(ipoa_A, 1, 1) = a
    phys_ocnatm(ipoa_seaice, 1, 1) = 0.0
    phys_ocnatm(ipoa_wspeed, 1, 1) = wspeed
    atm = 0.0
    atm(ia_pCO2) = pco2
    atm(ia_pO2) = po2
    call sub_calc_solconst(1, 1)
    call sub_calc_pv(1, 1)
    f = fun_calc_ocnatm_flux(1, 1, atm, dtyr)
    out(1) = ocnatm_airsea_solconst(ia_pCO2, 1, 1)
    out(2) = ocnatm_airsea_solconst(ia_pO2, 1, 1)
    out(3) = ocnatm_airsea_pv(ia_pCO2, 1, 1)
    out(4) = ocnatm_airsea_pv(ia_pO2, 1, 1)
    out(5) = f(ia_pCO2)
    out(6) = f(ia_pO2)
  end subroutine ref_airsea_cell

end module ref_airsea_shim
