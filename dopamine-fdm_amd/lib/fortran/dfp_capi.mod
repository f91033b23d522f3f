﻿!mod$ v1 sum:a82f29ac822f5217
!need$ 0bde2ac47243ead2 i iso_c_binding
!need$ bb381bf46e508468 i __fortran_builtins
module dfp_capi
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
use,intrinsic::__fortran_builtins,only:iso_c_binding$__fortran_builtins$c_associated_c_ptr=>c_associated_c_ptr
private::iso_c_binding$__fortran_builtins$c_associated_c_ptr
integer(4),parameter::dfp_fp64=0_4
integer(4),parameter::dfp_fp32_refine=1_4
integer(4),parameter::dfp_fields_on_device=1_4
integer(4),parameter::dfp_save_p=2_4
integer(4),parameter::dfp_async=4_4
type,bind(c)::dfp_grid_t
integer(4)::nxp
integer(4)::nym
integer(4)::nzp
real(8)::dx
real(8)::dz
type(c_ptr)::x
type(c_ptr)::y
type(c_ptr)::z
type(c_ptr)::yg
type(c_ptr)::zg
integer(4)::x_bc_type
integer(4)::y_bc_type
integer(4)::z_bc_type
end type
interface
function dfp_plan_create(g,precision,device,plan) bind(c,name="dfp_plan_create")
import::c_ptr
import::dfp_grid_t
type(dfp_grid_t),intent(in)::g
integer(4),value::precision
integer(4),value::device
type(c_ptr),intent(out)::plan
integer(4)::dfp_plan_create
end
end interface
interface
function dfp_plan_destroy(plan) bind(c,name="dfp_plan_destroy")
import::c_ptr
type(c_ptr),value::plan
integer(4)::dfp_plan_destroy
end
end interface
interface
function dfp_project(plan,u,v,w,rhs_p,p,dt,c,flags) bind(c,name="dfp_project")
import::c_ptr
type(c_ptr),value::plan
real(8),intent(inout)::u(1_8:*)
real(8),intent(inout)::v(1_8:*)
real(8),intent(inout)::w(1_8:*)
real(8),intent(inout)::rhs_p(1_8:*)
real(8),intent(inout)::p(1_8:*)
real(8),value::dt
real(8),value::c
integer(4),value::flags
integer(4)::dfp_project
end
end interface
interface
function dfp_divergence(plan,u,v,w,rhs_p,flags) bind(c,name="dfp_divergence")
import::c_ptr
type(c_ptr),value::plan
real(8),intent(in)::u(1_8:*)
real(8),intent(in)::v(1_8:*)
real(8),intent(in)::w(1_8:*)
real(8),intent(inout)::rhs_p(1_8:*)
integer(4),value::flags
integer(4)::dfp_divergence
end
end interface
interface
function dfp_solve(plan,rhs_p,p,dt,c,flags) bind(c,name="dfp_solve")
import::c_ptr
type(c_ptr),value::plan
real(8),intent(inout)::rhs_p(1_8:*)
real(8),intent(inout)::p(1_8:*)
real(8),value::dt
real(8),value::c
integer(4),value::flags
integer(4)::dfp_solve
end
end interface
interface
function dfp_correct(plan,u,v,w,rhs_p,flags) bind(c,name="dfp_correct")
import::c_ptr
type(c_ptr),value::plan
real(8),intent(inout)::u(1_8:*)
real(8),intent(inout)::v(1_8:*)
real(8),intent(inout)::w(1_8:*)
real(8),intent(in)::rhs_p(1_8:*)
integer(4),value::flags
integer(4)::dfp_correct
end
end interface
interface
function dfp_max_divergence(plan,u,v,w,out,flags) bind(c,name="dfp_max_divergence")
import::c_ptr
type(c_ptr),value::plan
real(8),intent(in)::u(1_8:*)
real(8),intent(in)::v(1_8:*)
real(8),intent(in)::w(1_8:*)
real(8),intent(out)::out
integer(4),value::flags
integer(4)::dfp_max_divergence
end
end interface
interface
function dfp_strerror(code) bind(c,name="dfp_strerror")
import::c_ptr
integer(4),value::code
type(c_ptr)::dfp_strerror
end
end interface
interface
function dfp_last_error(plan) bind(c,name="dfp_last_error")
import::c_ptr
type(c_ptr),value::plan
type(c_ptr)::dfp_last_error
end
end interface
contains
function c_string(p) result(s)
type(c_ptr),intent(in)::p
character(:,1),allocatable::s
end
end
