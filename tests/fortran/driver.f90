!> Test driver: reads a case written by tests/test_fortran_dropin.py, fills the global arrays with the
!  reference's lower bounds, runs the drop-in projection as time_integration does at RK stage 3, and
!  writes the results back (raw little-endian float64 streams).  The case gives the x, y and z boundary types
!  and, for the inflow/outflow x, the outlet backflow state (beta, outflow_Uc, Pb).
Program driver
  Use global
  Use projection
  Implicit None
  Integer :: u_in, u_out, n(6), nxp, nym, nzp
  Real(Int64) :: hdr(5), maxdiv
  Character(256) :: fin, fout

  Call get_command_argument(1, fin)
  Call get_command_argument(2, fout)
  Open(newunit=u_in, file=Trim(fin), access='stream', form='unformatted', status='old')
  Read(u_in) n
  nxp = n(1); nym = n(2); nzp = n(3); x_bc_type = n(4); y_bc_type = n(5); z_bc_type = n(6)
  nx = nxp + 2
  If ( x_bc_type == 1 ) nx = nxp + 1
  ny = nym + 1;  nz = nzp + 2
  If ( z_bc_type == 1 ) nz = nzp + 1
  nxg = nx + 1; nyg = nym + 2; nzg = nz + 1
  Read(u_in) hdr
  dx = hdr(1);  dz = hdr(2);  dt = hdr(3)
  outflow_backflow_beta = hdr(4);  outflow_Uc = hdr(5)
  Allocate( x(nx), y(ny), z(nz), yg(nyg), zg(nzg) )
  Read(u_in) x, y, z, yg, zg
  Allocate( U(nx,nyg,nzg), V(nxg,ny,nzg), W(nxg,nyg,nz), P(nxg,nyg,nzg) )
  Allocate( rhs_p(2:nxg, 2:nyg-1, 2:nzg) )
  Allocate( Pb(2:nyg-1, 2:nzg-1), phib(2:nyg-1, 2:nzg-1) )
  Read(u_in) U, V, W, Pb
  Close(u_in)
  P = 0d0
  rhs_p = 0d0
  phib = 0d0
  Allocate( rk_coef(3,3) )
  rk_coef = 0d0
  rk_coef(1,1) = 8d0/15d0;  rk_coef(2,2) = 5d0/12d0;  rk_coef(3,3) = 3d0/4d0

  rk_step = 3
  Call compute_projection_step
  Call check_divergence(maxdiv)

  Open(newunit=u_out, file=Trim(fout), access='stream', form='unformatted', status='replace')
  Write(u_out) U, V, W, rhs_p, P, maxdiv, phib
  Close(u_out)
  Call projection_release
End Program driver
