﻿!mod$ v1 sum:19e1f046a64d0c77
!need$ 9977815350a25db5 n global
!need$ a82f29ac822f5217 n dfp_capi
!need$ f1de5abe9bfe2168 i iso_fortran_env
!need$ 0bde2ac47243ead2 i iso_c_binding
!need$ bb381bf46e508468 i __fortran_builtins
module projection
use,intrinsic::iso_fortran_env,only:error_unit
use,intrinsic::iso_fortran_env,only:int32
use,intrinsic::iso_fortran_env,only:int64
use,intrinsic::iso_c_binding,only:c_associated
use,intrinsic::iso_c_binding,only:c_funloc
use,intrinsic::iso_c_binding,only:c_funptr
use,intrinsic::iso_c_binding,only:c_f_pointer
use,intrinsic::iso_c_binding,only:c_loc
use,intrinsic::iso_c_binding,only:c_null_funptr
use,intrinsic::iso_c_binding,only:c_null_ptr
use,intrinsic::iso_c_binding,only:c_ptr
use,intrinsic::iso_c_binding,only:c_sizeof
use,intrinsic::iso_c_binding,only:operator(==)
use,intrinsic::iso_c_binding,only:operator(/=)
use,intrinsic::iso_c_binding,only:c_int8_t
use,intrinsic::iso_c_binding,only:c_int16_t
use,intrinsic::iso_c_binding,only:c_int32_t
use,intrinsic::iso_c_binding,only:c_int64_t
use,intrinsic::iso_c_binding,only:c_int128_t
use,intrinsic::iso_c_binding,only:c_int
use,intrinsic::iso_c_binding,only:c_short
use,intrinsic::iso_c_binding,only:c_long
use,intrinsic::iso_c_binding,only:c_long_long
use,intrinsic::iso_c_binding,only:c_signed_char
use,intrinsic::iso_c_binding,only:c_size_t
use,intrinsic::iso_c_binding,only:c_intmax_t
use,intrinsic::iso_c_binding,only:c_intptr_t
use,intrinsic::iso_c_binding,only:c_ptrdiff_t
use,intrinsic::iso_c_binding,only:c_int_least8_t
use,intrinsic::iso_c_binding,only:c_int_fast8_t
use,intrinsic::iso_c_binding,only:c_int_least16_t
use,intrinsic::iso_c_binding,only:c_int_fast16_t
use,intrinsic::iso_c_binding,only:c_int_least32_t
use,intrinsic::iso_c_binding,only:c_int_fast32_t
use,intrinsic::iso_c_binding,only:c_int_least64_t
use,intrinsic::iso_c_binding,only:c_int_fast64_t
use,intrinsic::iso_c_binding,only:c_int_least128_t
use,intrinsic::iso_c_binding,only:c_int_fast128_t
use,intrinsic::iso_c_binding,only:c_float
use,intrinsic::iso_c_binding,only:c_double
use,intrinsic::iso_c_binding,only:c_long_double
use,intrinsic::iso_c_binding,only:c_float_complex
use,intrinsic::iso_c_binding,only:c_double_complex
use,intrinsic::iso_c_binding,only:c_long_double_complex
use,intrinsic::iso_c_binding,only:c_bool
use,intrinsic::iso_c_binding,only:c_char
use,intrinsic::iso_c_binding,only:c_null_char
use,intrinsic::iso_c_binding,only:c_alert
use,intrinsic::iso_c_binding,only:c_backspace
use,intrinsic::iso_c_binding,only:c_form_feed
use,intrinsic::iso_c_binding,only:c_new_line
use,intrinsic::iso_c_binding,only:c_carriage_return
use,intrinsic::iso_c_binding,only:c_horizontal_tab
use,intrinsic::iso_c_binding,only:c_vertical_tab
use,intrinsic::iso_c_binding,only:c_float128
use,intrinsic::iso_c_binding,only:c_float128_complex
use,intrinsic::iso_c_binding,only:c_uint8_t
use,intrinsic::iso_c_binding,only:c_uint16_t
use,intrinsic::iso_c_binding,only:c_uint32_t
use,intrinsic::iso_c_binding,only:c_uint64_t
use,intrinsic::iso_c_binding,only:c_uint128_t
use,intrinsic::iso_c_binding,only:c_unsigned_char
use,intrinsic::iso_c_binding,only:c_unsigned_short
use,intrinsic::iso_c_binding,only:c_unsigned
use,intrinsic::iso_c_binding,only:c_unsigned_long
use,intrinsic::iso_c_binding,only:c_unsigned_long_long
use,intrinsic::iso_c_binding,only:c_uintmax_t
use,intrinsic::iso_c_binding,only:c_uint_fast8_t
use,intrinsic::iso_c_binding,only:c_uint_fast16_t
use,intrinsic::iso_c_binding,only:c_uint_fast32_t
use,intrinsic::iso_c_binding,only:c_uint_fast64_t
use,intrinsic::iso_c_binding,only:c_uint_fast128_t
use,intrinsic::iso_c_binding,only:c_uint_least8_t
use,intrinsic::iso_c_binding,only:c_uint_least16_t
use,intrinsic::iso_c_binding,only:c_uint_least32_t
use,intrinsic::iso_c_binding,only:c_uint_least64_t
use,intrinsic::iso_c_binding,only:c_uint_least128_t
use,intrinsic::iso_c_binding,only:c_f_procpointer
use global,only:nx
use global,only:ny
use global,only:nz
use global,only:nxg
use global,only:nyg
use global,only:nzg
use global,only:x_bc_type
use global,only:y_bc_type
use global,only:z_bc_type
use global,only:rk_step
use global,only:dx
use global,only:dz
use global,only:dt
use global,only:x
use global,only:y
use global,only:z
use global,only:yg
use global,only:zg
use global,only:u
use global,only:v
use global,only:w
use global,only:p
use global,only:rhs_p
use global,only:rk_coef
use dfp_capi,only:dfp_fp64
use dfp_capi,only:dfp_fp32_refine
use dfp_capi,only:dfp_fields_on_device
use dfp_capi,only:dfp_save_p
use dfp_capi,only:dfp_async
use dfp_capi,only:dfp_grid_t
use dfp_capi,only:dfp_plan_create
use dfp_capi,only:dfp_plan_destroy
use dfp_capi,only:dfp_project
use dfp_capi,only:dfp_divergence
use dfp_capi,only:dfp_solve
use dfp_capi,only:dfp_correct
use dfp_capi,only:dfp_max_divergence
use dfp_capi,only:dfp_strerror
use dfp_capi,only:dfp_last_error
use dfp_capi,only:c_string
use,intrinsic::__fortran_builtins,only:iso_c_binding$__fortran_builtins$c_associated_c_ptr=>c_associated_c_ptr
private::error_unit
private::int32
private::int64
private::c_associated
private::c_funloc
private::c_funptr
private::c_f_pointer
private::c_loc
private::c_null_funptr
private::c_null_ptr
private::c_ptr
private::c_sizeof
private::operator(==)
private::operator(/=)
private::c_int8_t
private::c_int16_t
private::c_int32_t
private::c_int64_t
private::c_int128_t
private::c_int
private::c_short
private::c_long
private::c_long_long
private::c_signed_char
private::c_size_t
private::c_intmax_t
private::c_intptr_t
private::c_ptrdiff_t
private::c_int_least8_t
private::c_int_fast8_t
private::c_int_least16_t
private::c_int_fast16_t
private::c_int_least32_t
private::c_int_fast32_t
private::c_int_least64_t
private::c_int_fast64_t
private::c_int_least128_t
private::c_int_fast128_t
private::c_float
private::c_double
private::c_long_double
private::c_float_complex
private::c_double_complex
private::c_long_double_complex
private::c_bool
private::c_char
private::c_null_char
private::c_alert
private::c_backspace
private::c_form_feed
private::c_new_line
private::c_carriage_return
private::c_horizontal_tab
private::c_vertical_tab
private::c_float128
private::c_float128_complex
private::c_uint8_t
private::c_uint16_t
private::c_uint32_t
private::c_uint64_t
private::c_uint128_t
private::c_unsigned_char
private::c_unsigned_short
private::c_unsigned
private::c_unsigned_long
private::c_unsigned_long_long
private::c_uintmax_t
private::c_uint_fast8_t
private::c_uint_fast16_t
private::c_uint_fast32_t
private::c_uint_fast64_t
private::c_uint_fast128_t
private::c_uint_least8_t
private::c_uint_least16_t
private::c_uint_least32_t
private::c_uint_least64_t
private::c_uint_least128_t
private::c_f_procpointer
private::nx
private::ny
private::nz
private::nxg
private::nyg
private::nzg
private::x_bc_type
private::y_bc_type
private::z_bc_type
private::rk_step
private::dx
private::dz
private::dt
private::x
private::y
private::z
private::yg
private::zg
private::u
private::v
private::w
private::p
private::rhs_p
private::rk_coef
private::dfp_fp64
private::dfp_fp32_refine
private::dfp_fields_on_device
private::dfp_save_p
private::dfp_async
private::dfp_grid_t
private::dfp_plan_create
private::dfp_plan_destroy
private::dfp_project
private::dfp_divergence
private::dfp_solve
private::dfp_correct
private::dfp_max_divergence
private::dfp_strerror
private::dfp_last_error
private::c_string
private::iso_c_binding$__fortran_builtins$c_associated_c_ptr
type(c_ptr),private,save::plan
real(8),allocatable,private,save,target::gx(:)
real(8),allocatable,private,save,target::gy(:)
real(8),allocatable,private,save,target::gz(:)
real(8),allocatable,private,save,target::gyg(:)
real(8),allocatable,private,save,target::gzg(:)
real(8),private,save::no_p(1_8:1_8)
private::dfp_check
private::ensure_plan
contains
subroutine dfp_check(rc,where)
integer(4),intent(in)::rc
character(*,1),intent(in)::where
end
subroutine ensure_plan()
end
subroutine projection_release()
end
subroutine compute_projection_step()
end
subroutine compute_pseudo_pressure_rhs()
end
subroutine solve_poisson_equation(skip_p_save)
logical(4),intent(in),optional::skip_p_save
end
subroutine project_velocity()
end
subroutine apply_periodic_xz_pressure()
end
subroutine update_ghost_interior_planes_pressure()
end
subroutine update_ghost_interior_planes_pressure_x()
end
subroutine finish_pressure_ghosts(pf)
real(8),intent(inout)::pf(1_8:int(nxg,kind=8),1_8:int(nyg,kind=8),1_8:int(nzg,kind=8))
end
subroutine check_divergence(max_divergence)
real(8),intent(out)::max_divergence
end
end
