! 2-D PDE driver (the reference's tests/pdegen/mld_d_pde2d.f90).
program mld_d_pde2d
  use mld_pde_driver
  implicit none
  call pde_driver(2)
end program mld_d_pde2d
