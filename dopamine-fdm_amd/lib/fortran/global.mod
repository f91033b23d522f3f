﻿!mod$ v1 sum:9977815350a25db5
!need$ f1de5abe9bfe2168 i iso_fortran_env
module global
use,intrinsic::iso_fortran_env,only:int32
use,intrinsic::iso_fortran_env,only:int64
integer(4)::nx
integer(4)::ny
integer(4)::nz
integer(4)::nxg
integer(4)::nyg
integer(4)::nzg
integer(4)::x_bc_type
integer(4)::y_bc_type
integer(4)::z_bc_type
integer(4)::rk_step
real(8)::dx
real(8)::dz
real(8)::dt
real(8),allocatable::x(:)
real(8),allocatable::y(:)
real(8),allocatable::z(:)
real(8),allocatable::yg(:)
real(8),allocatable::zg(:)
real(8),allocatable::u(:,:,:)
real(8),allocatable::v(:,:,:)
real(8),allocatable::w(:,:,:)
real(8),allocatable::p(:,:,:)
real(8),allocatable::rhs_p(:,:,:)
real(8),allocatable::rk_coef(:,:)
end
