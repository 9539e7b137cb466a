! cgo_api_layout.f90 -- prints the sizes of the bind(C) types of cgo_api so
! that a test can compare them with the C structs of include/cgocn.h.
program cgo_api_layout
  use, intrinsic :: iso_c_binding
  use cgo_api
  implicit none
  type(cgo_grid_desc) :: g
  type(cgo_member_params) :: m
  type(cgo_bgc_params) :: b
  type(cgo_diag) :: d
  type(cgo_momentum_desc) :: md
  type(cgo_surflux_desc) :: sd
  write (*, '(a,i0)') 'cgo_grid_desc ', c_sizeof(g)
  write (*, '(a,i0)') 'cgo_member_params ', c_sizeof(m)
  write (*, '(a,i0)') 'cgo_bgc_params ', c_sizeof(b)
  write (*, '(a,i0)') 'cgo_diag ', c_sizeof(d)
  write (*, '(a,i0)') 'cgo_momentum_desc ', c_sizeof(md)
  write (*, '(a,i0)') 'cgo_surflux_desc ', c_sizeof(sd)
end program cgo_api_layout
