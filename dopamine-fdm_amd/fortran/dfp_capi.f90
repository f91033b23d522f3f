!> ISO_C_BINDING interface to the dfp C ABI (include/dfp.h).
Module dfp_capi

  Use iso_c_binding
  Implicit None

  Integer(c_int), Parameter :: DFP_FP64 = 0, DFP_FP32_REFINE = 1
  Integer(c_int), Parameter :: DFP_FIELDS_ON_DEVICE = 1, DFP_SAVE_P = 2, DFP_ASYNC = 4

  !> mirrors struct dfp_grid
  Type, Bind(C) :: dfp_grid_t
     Integer(c_int32_t) :: nxp, nym, nzp
     Real(c_double)     :: dx, dz
     Type(c_ptr)        :: x, y, z, yg, zg
     Integer(c_int32_t) :: x_bc_type, y_bc_type, z_bc_type
  End Type dfp_grid_t

  Interface
     Integer(c_int) Function dfp_plan_create(g, precision, device, plan) Bind(C, name='dfp_plan_create')
       Import :: c_int, c_ptr, dfp_grid_t
       Type(dfp_grid_t), Intent(In) :: g
       Integer(c_int), Value :: precision, device
       Type(c_ptr), Intent(Out) :: plan
     End Function dfp_plan_create

     Integer(c_int) Function dfp_plan_destroy(plan) Bind(C, name='dfp_plan_destroy')
       Import :: c_int, c_ptr
       Type(c_ptr), Value :: plan
     End Function dfp_plan_destroy

     Integer(c_int) Function dfp_project(plan, U, V, W, rhs_p, P, dt, c, flags) Bind(C, name='dfp_project')
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: plan
       Real(c_double), Intent(InOut) :: U(*), V(*), W(*), rhs_p(*), P(*)
       Real(c_double), Value :: dt, c
       Integer(c_int), Value :: flags
     End Function dfp_project

     Integer(c_int) Function dfp_divergence(plan, U, V, W, rhs_p, flags) Bind(C, name='dfp_divergence')
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: plan
       Real(c_double), Intent(In) :: U(*), V(*), W(*)
       Real(c_double), Intent(InOut) :: rhs_p(*)
       Integer(c_int), Value :: flags
     End Function dfp_divergence

     Integer(c_int) Function dfp_solve(plan, rhs_p, P, dt, c, flags) Bind(C, name='dfp_solve')
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: plan
       Real(c_double), Intent(InOut) :: rhs_p(*), P(*)
       Real(c_double), Value :: dt, c
       Integer(c_int), Value :: flags
     End Function dfp_solve

     Integer(c_int) Function dfp_correct(plan, U, V, W, rhs_p, flags) Bind(C, name='dfp_correct')
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: plan
       Real(c_double), Intent(InOut) :: U(*), V(*), W(*)
       Real(c_double), Intent(In) :: rhs_p(*)
       Integer(c_int), Value :: flags
     End Function dfp_correct

     Integer(c_int) Function dfp_set_outflow(plan, beta, outflow_Uc, Pb, phib) Bind(C, name='dfp_set_outflow')
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: plan
       Real(c_double), Value :: beta, outflow_Uc
       Real(c_double), Intent(In) :: Pb(*)
       Real(c_double), Intent(InOut) :: phib(*)
     End Function dfp_set_outflow

     Integer(c_int) Function dfp_outflow_backflow(plan, U, V, W, rhs_p, dt, c, flags) &
          Bind(C, name='dfp_outflow_backflow')
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: plan
       Real(c_double), Intent(In) :: U(*), V(*), W(*)
       Real(c_double), Intent(InOut) :: rhs_p(*)
       Real(c_double), Value :: dt, c
       Integer(c_int), Value :: flags
     End Function dfp_outflow_backflow

     Integer(c_int) Function dfp_max_divergence(plan, U, V, W, out, flags) Bind(C, name='dfp_max_divergence')
       Import :: c_int, c_ptr, c_double
       Type(c_ptr), Value :: plan
       Real(c_double), Intent(In) :: U(*), V(*), W(*)
       Real(c_double), Intent(Out) :: out
       Integer(c_int), Value :: flags
     End Function dfp_max_divergence

     Type(c_ptr) Function dfp_strerror(code) Bind(C, name='dfp_strerror')
       Import :: c_int, c_ptr
       Integer(c_int), Value :: code
     End Function dfp_strerror

     Type(c_ptr) Function dfp_last_error(plan) Bind(C, name='dfp_last_error')
       Import :: c_ptr
       Type(c_ptr), Value :: plan
     End Function dfp_last_error
  End Interface

Contains

  !> A C string as a Fortran character variable
  Function c_string(p) Result(s)
    Type(c_ptr), Intent(In) :: p
    Character(:), Allocatable :: s
    Character(kind=c_char), Pointer :: a(:)
    Integer :: n
    Interface
       Integer(c_size_t) Function strlen(q) Bind(C, name='strlen')
         Import :: c_size_t, c_ptr
         Type(c_ptr), Value :: q
       End Function strlen
    End Interface
    If ( .Not. c_associated(p) ) Then
       s = ''
       Return
    End If
    n = Int(strlen(p))
    Call c_f_pointer(p, a, [n])
    Allocate(Character(n) :: s)
    s = Transfer(a(1:n), s)
  End Function c_string

End Module dfp_capi
