!> Test stand-in for the part of the solver's global module (global.f90:130-197, :293-294) that the
!  drop-in projection module reads: the grid, the fields and the time-step data of one rank.
Module global
  Use iso_fortran_env, Only : Int32, Int64
  Implicit None
  Integer(Int32) :: nx, ny, nz, nxg, nyg, nzg
  Integer(Int32) :: x_bc_type = 0, y_bc_type = 1, z_bc_type = 0
  Integer(Int32) :: rk_step = 1
  Real(Int64) :: dx, dz, dt
  Real(Int64), Allocatable, Dimension(:) :: x, y, z, yg, zg
  ! TARGET: the test driver copies them to device memory it allocates itself
  Real(Int64), Allocatable, Dimension(:,:,:), Target :: U, V, W, P, rhs_p
  Real(Int64), Allocatable, Dimension(:,:) :: rk_coef
  ! outlet backflow stabilisation of the inflow/outflow x (global.f90:188-197)
  Real(Int64) :: outflow_backflow_beta = 0d0, outflow_Uc = 0d0
  Real(Int64), Allocatable, Dimension(:,:), Target :: Pb, phib
End Module global
