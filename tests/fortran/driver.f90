!> The HIP runtime calls the driver's device mode needs (ISO_C_BINDING to libamdhip64)
Module hip_min
  Use iso_c_binding
  Implicit None
  Integer(c_int), Parameter :: hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2
  Interface
     Integer(c_int) Function hipMalloc(p, n) Bind(C, name='hipMalloc')
       Import :: c_int, c_ptr, c_size_t
       Type(c_ptr) :: p
       Integer(c_size_t), Value :: n
     End Function hipMalloc
     Integer(c_int) Function hipFree(p) Bind(C, name='hipFree')
       Import :: c_int, c_ptr
       Type(c_ptr), Value :: p
     End Function hipFree
     Integer(c_int) Function hipMemcpy(dst, src, n, kind) Bind(C, name='hipMemcpy')
       Import :: c_int, c_ptr, c_size_t
       Type(c_ptr), Value :: dst, src
       Integer(c_size_t), Value :: n
       Integer(c_int), Value :: kind
     End Function hipMemcpy
  End Interface
Contains
  !> device copy of a host array of n doubles
  Function to_device(h, n) Result(d)
    Type(c_ptr), Intent(In) :: h
    Integer(c_size_t), Intent(In) :: n
    Type(c_ptr) :: d
    If ( hipMalloc(d, 8_c_size_t * n) /= 0 ) Error Stop 'hipMalloc'
    If ( hipMemcpy(d, h, 8_c_size_t * n, hipMemcpyHostToDevice) /= 0 ) Error Stop 'hipMemcpy'
  End Function to_device
  !> the device array back into the host array, then freed
  Subroutine from_device(h, d, n)
    Type(c_ptr), Intent(In) :: h, d
    Integer(c_size_t), Intent(In) :: n
    If ( hipMemcpy(h, d, 8_c_size_t * n, hipMemcpyDeviceToHost) /= 0 ) Error Stop 'hipMemcpy'
    If ( hipFree(d) /= 0 ) Error Stop 'hipFree'
  End Subroutine from_device
End Module hip_min

!> Test driver: reads a case written by tests/test_fortran_dropin.py, fills the global arrays with the
!  reference's lower bounds, runs the drop-in projection as time_integration does at RK stage 3, and
!  writes the results back (raw little-endian float64 streams).  The case gives the x, y and z boundary types
!  and, for the inflow/outflow x, the outlet backflow state (beta, outflow_Uc, Pb).  With a third argument
!  "device" the fields are first copied to device memory the driver allocates, and the device-resident
!  entry points run on them.  A fourth argument reps > 0 times the drop-in instead: after one untimed
!  call, reps calls at RK stages 1, 2, 3, 1, ... (P saved at stage 3), each returning with its results
!  in place, and prints "us_per_call <microseconds>"; no output file is written.
Program driver
  Use iso_c_binding
  Use global
  Use projection
  Use hip_min
  Implicit None
  Integer :: u_in, u_out, n(6), nxp, nym, nzp, reps, r
  Integer(Int64) :: c0, c1, crate
  Real(Int64) :: hdr(5), maxdiv
  Character(256) :: fin, fout, mode, arg
  Type(c_ptr) :: Ud, Vd, Wd, rd, Pd, Pbd, phd
  Logical :: dev

  Call get_command_argument(1, fin)
  Call get_command_argument(2, fout)
  Call get_command_argument(3, mode)
  dev = Trim(mode) == 'device'
  reps = 0
  If ( command_argument_count() >= 4 ) Then
     Call get_command_argument(4, arg)
     Read(arg, *) reps
  End If
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
  If ( dev ) Then
     Ud = to_device(c_loc(U), Size(U, kind=c_size_t));  Vd = to_device(c_loc(V), Size(V, kind=c_size_t))
     Wd = to_device(c_loc(W), Size(W, kind=c_size_t));  rd = to_device(c_loc(rhs_p), Size(rhs_p, kind=c_size_t))
     Pd = to_device(c_loc(P), Size(P, kind=c_size_t))
     Pbd = c_null_ptr;  phd = c_null_ptr
     If ( outflow_backflow_beta > 0d0 ) Then
        Pbd = to_device(c_loc(Pb), Size(Pb, kind=c_size_t));  phd = to_device(c_loc(phib), Size(phib, kind=c_size_t))
     End If
  End If
  If ( reps > 0 ) Then
     Call project_once
     Call system_clock(c0, crate)
     Do r = 1, reps
        rk_step = Mod(r - 1, 3) + 1
        Call project_once
     End Do
     Call system_clock(c1)
     Write(*, '(A, F12.2)') 'us_per_call ', 1d6 * Real(c1 - c0, Int64) / Real(crate, Int64) / reps
     Call projection_release
     Stop
  End If
  If ( dev ) Then
     Call compute_projection_step_device(Ud, Vd, Wd, rd, Pd, Pbd, phd)
     Call check_divergence_device(Ud, Vd, Wd, maxdiv)
     Call from_device(c_loc(U), Ud, Size(U, kind=c_size_t));  Call from_device(c_loc(V), Vd, Size(V, kind=c_size_t))
     Call from_device(c_loc(W), Wd, Size(W, kind=c_size_t))
     Call from_device(c_loc(rhs_p), rd, Size(rhs_p, kind=c_size_t));  Call from_device(c_loc(P), Pd, Size(P, kind=c_size_t))
     If ( outflow_backflow_beta > 0d0 ) Then
        Call from_device(c_loc(Pb), Pbd, Size(Pb, kind=c_size_t))
        Call from_device(c_loc(phib), phd, Size(phib, kind=c_size_t))
     End If
  Else
     Call compute_projection_step
     Call check_divergence(maxdiv)
  End If

  Open(newunit=u_out, file=Trim(fout), access='stream', form='unformatted', status='replace')
  Write(u_out) U, V, W, rhs_p, P, maxdiv, phib
  Close(u_out)
  Call projection_release

Contains

  !> one projection through the drop-in, host arrays or device copies
  Subroutine project_once
    If ( dev ) Then
       Call compute_projection_step_device(Ud, Vd, Wd, rd, Pd, Pbd, phd)
    Else
       Call compute_projection_step
    End If
  End Subroutine project_once

End Program driver
