!> Drop-in replacement of the reference's module projection (src/projection.f90) for z periodic with y walls
!  and a periodic or an inflow/outflow x (x_bc_type 0 or 1, with the outlet backflow term), or with y and x
!  periodic (y_bc_type 0, the periodic box), on one rank: the
!  same subroutine names and arguments, backed by the dfp HIP library through its C ABI.  The fields are the
!  global-module arrays (global.f90:139-146, :188-197); they are copied to the device and back on each call.
!  A caller that keeps them on the device calls compute_projection_step_device with their device addresses
!  instead (no copies).
Module projection

  Use iso_c_binding
  Use iso_fortran_env, Only : error_unit, Int32, Int64
  Use global
  Use dfp_capi

  Implicit None

  Private
  Public :: compute_projection_step, compute_pseudo_pressure_rhs, set_outflow_backflow_pressure, &
            solve_poisson_equation, project_velocity, apply_periodic_xz_pressure, &
            update_ghost_interior_planes_pressure, update_ghost_interior_planes_pressure_x, &
            finish_pressure_ghosts, check_divergence, projection_release, &
            compute_projection_step_device, check_divergence_device

  Type(c_ptr), Save :: plan = c_null_ptr
  ! the grid the plan was built from (c_loc needs TARGET arrays)
  Real(c_double), Allocatable, Target, Save :: gx(:), gy(:), gz(:), gyg(:), gzg(:)
  ! stands in for P when no pressure is saved, and for Pb and phib without the backflow term
  Real(c_double), Save :: no_p(1) = 0d0
  Real(c_double), Save :: no_face(1) = 0d0

Contains

  !> Stops with the library's message (the reference aborts through gpu_abort, poisson_gpu.f90:51-56)
  Subroutine dfp_check(rc, where)
    Integer(c_int), Intent(In) :: rc
    Character(*), Intent(In) :: where
    If ( rc /= 0 ) Then
       Write(error_unit,'(A)') 'ERROR: '//where//': '//c_string(dfp_strerror(rc))//': '//c_string(dfp_last_error(plan))
       Error Stop 1
    End If
  End Subroutine dfp_check

  !> Builds the plan on first use from the global grid (the reference builds its GPU plans lazily,
  !  poisson_gpu.f90:234, :436)
  Subroutine ensure_plan
    Type(dfp_grid_t) :: g
    If ( c_associated(plan) ) Return
    ! the channel (either x type), the periodic box (y_bc_type = 0 with a periodic x, initialization.f90:96-97)
    ! or the 4-wall duct (z_bc_type = 1 with a periodic x and y walls, :98-100)
    If ( ( x_bc_type /= 0 .And. x_bc_type /= 1 ) .Or. ( y_bc_type /= 0 .And. y_bc_type /= 1 ) .Or. &
         ( y_bc_type == 0 .And. x_bc_type /= 0 ) .Or. ( z_bc_type /= 0 .And. z_bc_type /= 1 ) .Or. &
         ( z_bc_type == 1 .And. ( x_bc_type /= 0 .Or. y_bc_type /= 1 ) ) ) Then
       Write(error_unit,'(A)') 'ERROR: the dfp projection supports the channel (x periodic or inflow/outflow), ' // &
            'the periodic box and the 4-wall duct only'
       Error Stop 1
    End If
    gx = x;  gy = y;  gz = z;  gyg = yg;  gzg = zg
    ! the periodic x drops its duplicate cell, the inflow/outflow x keeps all nxm (initialization.f90:331-338)
    g%nxp = nx - 2
    If ( x_bc_type == 1 ) g%nxp = nx - 1
    ! z walls keep all nzm cells too (initialization.f90:339-345)
    g%nym = ny - 1;  g%nzp = nz - 2
    If ( z_bc_type == 1 ) g%nzp = nz - 1
    g%dx = dx;  g%dz = dz
    g%x = c_loc(gx);  g%y = c_loc(gy);  g%z = c_loc(gz);  g%yg = c_loc(gyg);  g%zg = c_loc(gzg)
    g%x_bc_type = x_bc_type;  g%y_bc_type = y_bc_type;  g%z_bc_type = z_bc_type
    Call dfp_check(dfp_plan_create(g, DFP_FP64, 0_c_int, plan), 'dfp_plan_create')
  End Subroutine ensure_plan

  !> Releases the plan (no counterpart in the reference; call at finalisation)
  Subroutine projection_release
    Integer(c_int) :: rc
    If ( c_associated(plan) ) rc = dfp_plan_destroy(plan)
    plan = c_null_ptr
  End Subroutine projection_release

  !> Hands the outlet backflow state (global.f90:188-197) to the plan before a call that reads it: the plan
  !  reads Pb and writes phib during that call (outflow_Uc changes every step, boundary_conditions.f90:208)
  Subroutine sync_outflow
    If ( outflow_backflow_beta > 0d0 ) Then
       Call dfp_check(dfp_set_outflow(plan, outflow_backflow_beta, outflow_Uc, Pb, phib), 'dfp_set_outflow')
    Else
       Call dfp_check(dfp_set_outflow(plan, 0d0, outflow_Uc, no_face, no_face), 'dfp_set_outflow')
    End If
  End Subroutine sync_outflow

  !> projection.f90:21-34 (the backflow term of :24 included); c = rk_coef(rk_step,rk_step), which at the last
  !  stage is the rk_coef(3,3) of the pressure save
  Subroutine compute_projection_step
    Integer(c_int) :: flags
    Call ensure_plan
    Call sync_outflow
    flags = 0
    If ( rk_step == 3 ) flags = DFP_SAVE_P
    If ( rk_step == 3 ) Then
       Call dfp_check(dfp_project(plan, U, V, W, rhs_p, P, dt, rk_coef(rk_step,rk_step), flags), 'dfp_project')
    Else
       Call dfp_check(dfp_project(plan, U, V, W, rhs_p, no_p, dt, rk_coef(rk_step,rk_step), flags), 'dfp_project')
    End If
  End Subroutine compute_projection_step

  !> projection.f90:65-96
  Subroutine set_outflow_backflow_pressure
    Call ensure_plan
    Call sync_outflow
    Call dfp_check(dfp_outflow_backflow(plan, U, V, W, rhs_p, dt, rk_coef(rk_step,rk_step), 0_c_int), &
         'dfp_outflow_backflow')
  End Subroutine set_outflow_backflow_pressure

  !> projection.f90:37-63
  Subroutine compute_pseudo_pressure_rhs
    Call ensure_plan
    Call dfp_check(dfp_divergence(plan, U, V, W, rhs_p, 0_c_int), 'dfp_divergence')
  End Subroutine compute_pseudo_pressure_rhs

  !> projection.f90:100-382
  Subroutine solve_poisson_equation(skip_p_save)
    Logical, Intent(In), Optional :: skip_p_save
    Logical :: save_p
    save_p = .True.
    If ( Present(skip_p_save) ) save_p = .Not. skip_p_save
    Call ensure_plan
    If ( rk_step == 3 .And. save_p ) Then
       Call dfp_check(dfp_solve(plan, rhs_p, P, dt, rk_coef(3,3), DFP_SAVE_P), 'dfp_solve')
    Else
       Call dfp_check(dfp_solve(plan, rhs_p, no_p, dt, rk_coef(3,3), 0_c_int), 'dfp_solve')
    End If
  End Subroutine solve_poisson_equation

  !> projection.f90:551-604 (the outlet face with phib, :578-589)
  Subroutine project_velocity
    Call ensure_plan
    Call sync_outflow
    Call dfp_check(dfp_correct(plan, U, V, W, rhs_p, 0_c_int), 'dfp_correct')
  End Subroutine project_velocity

  !> projection.f90:441-500 on one rank: the periodic duplicate planes of rhs_p (the inflow/outflow x has
  !  none in x, :446-448)
  Subroutine apply_periodic_xz_pressure
    If ( x_bc_type == 0 ) rhs_p ( nxg-1, :, : ) = rhs_p ( 2, :, : )
    rhs_p ( 2:nxg-1, :, nz ) = rhs_p ( 2:nxg-1, :, 2 )
  End Subroutine apply_periodic_xz_pressure

  !> projection.f90:503-524: on one rank there is no z neighbour, nothing to exchange
  Subroutine update_ghost_interior_planes_pressure
  End Subroutine update_ghost_interior_planes_pressure

  !> projection.f90:529-548: on one rank there is no x neighbour, nothing to exchange
  Subroutine update_ghost_interior_planes_pressure_x
  End Subroutine update_ghost_interior_planes_pressure_x

  !> projection.f90:388-438 with boundary_conditions.f90:277-289, :591-602 (id = 4) on one rank; Neumann
  !  copies at the two ends of the inflow/outflow x (:413-422)
  Subroutine finish_pressure_ghosts(Pf)
    Real(Int64), Intent(InOut) :: Pf(nxg,nyg,nzg)
    Pf(:,  1,:) = Pf(:,    2,:)
    Pf(:,nyg,:) = Pf(:,nyg-1,:)
    If ( x_bc_type == 0 ) Then
       Pf(    1,:,:) = Pf(nxg-2,:,:)
       Pf(nxg-1,:,:) = Pf(    2,:,:)
       Pf(nxg  ,:,:) = Pf(    3,:,:)
    Else
       Pf(  1,:,:) = Pf(    2,:,:)
       Pf(nxg,:,:) = Pf(nxg-1,:,:)
    End If
    Pf(:,:,    1) = Pf(:,:,nzg-2)
    Pf(:,:,nzg-1) = Pf(:,:,    2)
    Pf(:,:,nzg  ) = Pf(:,:,    3)
  End Subroutine finish_pressure_ghosts

  !> projection.f90:607-639 (no immersed boundary)
  Subroutine check_divergence(max_divergence)
    Real(Int64), Intent(Out) :: max_divergence
    Real(c_double) :: m
    Call ensure_plan
    Call dfp_check(dfp_max_divergence(plan, U, V, W, m, 0_c_int), 'dfp_max_divergence')
    max_divergence = m
  End Subroutine check_divergence

  !> A device address as an array argument of the C ABI: only the address travels (the pointer is
  !  contiguous, so no copy of the target is made on the host)
  Subroutine as_array(addr, a)
    Type(c_ptr), Intent(In) :: addr
    Real(c_double), Pointer, Contiguous, Intent(Out) :: a(:)
    Call c_f_pointer(addr, a, [1])
  End Subroutine as_array

  !> compute_projection_step (projection.f90:21-34) on fields the caller keeps on the device, as the
  !  reference's OpenACC build keeps U, V, W, rhs_p and P present there: their device addresses (for
  !  example from !$acc host_data use_device, or hipMalloc), with the global module's shapes; with the
  !  outlet backflow term also those of Pb and phib, else c_null_ptr.  The grid and the time-step state
  !  are the module's, as for compute_projection_step.  No copies over PCIe.
  Subroutine compute_projection_step_device(U_d, V_d, W_d, rhs_p_d, P_d, Pb_d, phib_d)
    Type(c_ptr), Intent(In) :: U_d, V_d, W_d, rhs_p_d, P_d, Pb_d, phib_d
    Real(c_double), Pointer, Contiguous :: ud(:), vd(:), wd(:), rd(:), pd(:), pbd(:), phd(:)
    Integer(c_int) :: flags
    Call ensure_plan
    Call as_array(U_d, ud);  Call as_array(V_d, vd);  Call as_array(W_d, wd)
    Call as_array(rhs_p_d, rd);  Call as_array(P_d, pd)
    If ( outflow_backflow_beta > 0d0 ) Then
       Call as_array(Pb_d, pbd);  Call as_array(phib_d, phd)
       Call dfp_check(dfp_set_outflow(plan, outflow_backflow_beta, outflow_Uc, pbd, phd), 'dfp_set_outflow')
    Else
       Call dfp_check(dfp_set_outflow(plan, 0d0, outflow_Uc, no_face, no_face), 'dfp_set_outflow')
    End If
    flags = DFP_FIELDS_ON_DEVICE
    If ( rk_step == 3 ) flags = flags + DFP_SAVE_P
    Call dfp_check(dfp_project(plan, ud, vd, wd, rd, pd, dt, rk_coef(rk_step,rk_step), flags), 'dfp_project')
  End Subroutine compute_projection_step_device

  !> check_divergence on device-resident U, V, W (their device addresses)
  Subroutine check_divergence_device(U_d, V_d, W_d, max_divergence)
    Type(c_ptr), Intent(In) :: U_d, V_d, W_d
    Real(Int64), Intent(Out) :: max_divergence
    Real(c_double), Pointer, Contiguous :: ud(:), vd(:), wd(:)
    Real(c_double) :: m
    Call ensure_plan
    Call as_array(U_d, ud);  Call as_array(V_d, vd);  Call as_array(W_d, wd)
    Call dfp_check(dfp_max_divergence(plan, ud, vd, wd, m, DFP_FIELDS_ON_DEVICE), 'dfp_max_divergence')
    max_divergence = m
  End Subroutine check_divergence_device

End Module projection
