! 3-D PDE driver (the reference's tests/pdegen/mld_d_pde3d.f90).
program mld_d_pde3d
  use mld_pde_driver
  implicit none
  call pde_driver(3)
end program mld_d_pde3d
