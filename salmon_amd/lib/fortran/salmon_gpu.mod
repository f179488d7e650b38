﻿!mod$ v1 sum:59e397f2545f5f00
!need$ 0bde2ac47243ead2 i iso_c_binding
!need$ bb381bf46e508468 i __fortran_builtins
module salmon_gpu
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
integer(4),parameter::salmon_ok=0_4
integer(4),parameter::salmon_err_arg=1_4
integer(4),parameter::salmon_err_hip=2_4
integer(4),parameter::salmon_err_state=3_4
integer(4),parameter::salmon_err_nomem=4_4
integer(4),parameter::salmon_err_unsupported=5_4
integer(4),parameter::salmon_fp64=0_4
integer(4),parameter::salmon_mixed=1_4
integer(4),parameter::salmon_path_auto=0_4
integer(4),parameter::salmon_path_streamed=1_4
integer(4),parameter::salmon_path_fused=2_4
integer(4),parameter::salmon_pc_none=0_4
integer(4),parameter::salmon_pc_kinetic=1_4
integer(4),parameter::salmon_pc_learned=2_4
type,bind(c)::salmon_grid_t
integer(4)::nlx
integer(4)::nly
integer(4)::nlz
real(8)::alx
real(8)::aly
real(8)::alz
real(8)::lapx(0_8:4_8)
real(8)::lapy(0_8:4_8)
real(8)::lapz(0_8:4_8)
real(8)::nabx(0_8:4_8)
real(8)::naby(0_8:4_8)
real(8)::nabz(0_8:4_8)
end type
interface
function salmon_gpu_init(grid,precision,ctx) bind(c,name="salmon_gpu_init")
import::c_ptr
import::salmon_grid_t
type(salmon_grid_t),intent(in)::grid
integer(4),value::precision
type(c_ptr),intent(out)::ctx
integer(4)::salmon_gpu_init
end
end interface
interface
function salmon_gpu_set_potential(ctx,vloc) bind(c,name="salmon_gpu_set_potential")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(in)::vloc(1_8:*)
integer(4)::salmon_gpu_set_potential
end
end interface
interface
function salmon_gpu_set_projectors(ctx,ni,nps,mps,jxyz,jxx,jyy,jzz,nlma,a_tbl,uv,iuv) bind(c,name="salmon_gpu_set_projectors")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::ni
integer(4),value::nps
integer(4),intent(in)::mps(1_8:*)
integer(4),intent(in)::jxyz(1_8:*)
integer(4),intent(in)::jxx(1_8:*)
integer(4),intent(in)::jyy(1_8:*)
integer(4),intent(in)::jzz(1_8:*)
integer(4),value::nlma
integer(4),intent(in)::a_tbl(1_8:*)
real(8),intent(in)::uv(1_8:*)
integer(4),intent(in)::iuv(1_8:*)
integer(4)::salmon_gpu_set_projectors
end
end interface
interface
function salmon_gpu_set_kpoints(ctx,nk,nb,kac0,occ) bind(c,name="salmon_gpu_set_kpoints")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::nk
integer(4),value::nb
real(8),intent(in)::kac0(1_8:*)
real(8),intent(in)::occ(1_8:*)
integer(4)::salmon_gpu_set_kpoints
end
end interface
interface
function salmon_gpu_put_orbitals(ctx,zu) bind(c,name="salmon_gpu_put_orbitals")
import::c_ptr
type(c_ptr),value::ctx
complex(8),intent(in)::zu(1_8:*)
integer(4)::salmon_gpu_put_orbitals
end
end interface
interface
function salmon_gpu_get_orbitals(ctx,zu) bind(c,name="salmon_gpu_get_orbitals")
import::c_ptr
type(c_ptr),value::ctx
complex(8),intent(out)::zu(1_8:*)
integer(4)::salmon_gpu_get_orbitals
end
end interface
interface
function salmon_gpu_hpsi(ctx,kac,tpsi,htpsi,ttpsi) bind(c,name="salmon_gpu_hpsi")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(in)::kac(1_8:3_8)
type(c_ptr),value::tpsi
type(c_ptr),value::htpsi
type(c_ptr),value::ttpsi
integer(4)::salmon_gpu_hpsi
end
end interface
interface
function salmon_gpu_hamiltonian(ctx,kac,dt) bind(c,name="salmon_gpu_hamiltonian")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(in)::kac(1_8:*)
real(8),value::dt
integer(4)::salmon_gpu_hamiltonian
end
end interface
interface
function salmon_gpu_current(ctx,kac,jav) bind(c,name="salmon_gpu_current")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(in)::kac(1_8:*)
real(8),intent(out)::jav(1_8:3_8)
integer(4)::salmon_gpu_current
end
end interface
interface
function salmon_gpu_density(ctx,rho) bind(c,name="salmon_gpu_density")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(out)::rho(1_8:*)
integer(4)::salmon_gpu_density
end
end interface
interface
function salmon_gpu_energy(ctx,kac,ekin,enl,eloc) bind(c,name="salmon_gpu_energy")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(in)::kac(1_8:*)
real(8),intent(out)::ekin
real(8),intent(out)::enl
real(8),intent(out)::eloc
integer(4)::salmon_gpu_energy
end
end interface
interface
function salmon_gpu_norm(ctx,norm) bind(c,name="salmon_gpu_norm")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(out)::norm
integer(4)::salmon_gpu_norm
end
end interface
interface
function salmon_gpu_gram_schmidt(ctx) bind(c,name="salmon_gpu_gram_schmidt")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::salmon_gpu_gram_schmidt
end
end interface
interface
function salmon_gpu_sp_energy(ctx,kac,esp) bind(c,name="salmon_gpu_sp_energy")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(in)::kac(1_8:*)
real(8),intent(out)::esp(1_8:*)
integer(4)::salmon_gpu_sp_energy
end
end interface
interface
function salmon_gpu_cg(ctx,kac,iter_cg_max,precond,esp_var,iters) bind(c,name="salmon_gpu_cg")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(in)::kac(1_8:*)
integer(4),value::iter_cg_max
integer(4),value::precond
real(8),intent(out)::esp_var(1_8:*)
type(c_ptr),value::iters
integer(4)::salmon_gpu_cg
end
end interface
interface
function salmon_gpu_subspace_diag(ctx,kac,esp) bind(c,name="salmon_gpu_subspace_diag")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(in)::kac(1_8:*)
real(8),intent(out)::esp(1_8:*)
integer(4)::salmon_gpu_subspace_diag
end
end interface
interface
function salmon_gpu_ground_state(ctx,kac,max_sweeps,iter_cg_max,precond,tol,esp,esp_var,sweeps) bind(c,name="salmon_gpu_ground_state")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(in)::kac(1_8:*)
integer(4),value::max_sweeps
integer(4),value::iter_cg_max
integer(4),value::precond
real(8),value::tol
real(8),intent(out)::esp(1_8:*)
real(8),intent(out)::esp_var(1_8:*)
integer(4),intent(out)::sweeps
integer(4)::salmon_gpu_ground_state
end
end interface
interface
function salmon_gpu_rt_run(ctx,ac,nsteps,dt,reortho_every,jt,et,norm,rho) bind(c,name="salmon_gpu_rt_run")
import::c_ptr
type(c_ptr),value::ctx
real(8),intent(in)::ac(1_8:*)
integer(4),value::nsteps
real(8),value::dt
integer(4),value::reortho_every
type(c_ptr),value::jt
type(c_ptr),value::et
type(c_ptr),value::norm
type(c_ptr),value::rho
integer(4)::salmon_gpu_rt_run
end
end interface
interface
function salmon_gpu_set_path(ctx,path) bind(c,name="salmon_gpu_set_path")
import::c_ptr
type(c_ptr),value::ctx
integer(4),value::path
integer(4)::salmon_gpu_set_path
end
end interface
interface
function salmon_gpu_path(ctx) bind(c,name="salmon_gpu_path")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::salmon_gpu_path
end
end interface
interface
function salmon_gpu_synchronize(ctx) bind(c,name="salmon_gpu_synchronize")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::salmon_gpu_synchronize
end
end interface
interface
function salmon_gpu_stream(ctx) bind(c,name="salmon_gpu_stream")
import::c_ptr
type(c_ptr),value::ctx
type(c_ptr)::salmon_gpu_stream
end
end interface
interface
function salmon_gpu_errstr(code) bind(c,name="salmon_gpu_errstr")
import::c_ptr
integer(4),value::code
type(c_ptr)::salmon_gpu_errstr
end
end interface
interface
function salmon_gpu_last_error(ctx) bind(c,name="salmon_gpu_last_error")
import::c_ptr
type(c_ptr),value::ctx
type(c_ptr)::salmon_gpu_last_error
end
end interface
interface
function salmon_gpu_finalize(ctx) bind(c,name="salmon_gpu_finalize")
import::c_ptr
type(c_ptr),value::ctx
integer(4)::salmon_gpu_finalize
end
end interface
contains
function salmon_gpu_message(ctx,code) result(msg)
type(c_ptr),intent(in)::ctx
integer(4),intent(in)::code
character(:,1),allocatable::msg
end
function c_string(p) result(s)
type(c_ptr),intent(in)::p
character(:,1),allocatable::s
end
end
