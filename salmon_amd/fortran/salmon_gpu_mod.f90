! ISO_C_BINDING interface of the gfx950 ARTED hot path; one declaration per
! entry point of include/salmon_gpu.h (same names, same argument meaning).
! Arrays are passed as the Fortran arrays of the reference's module globals
! (column-major, complex(8) interleaved).
module salmon_gpu
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: SALMON_OK = 0
  integer(c_int), parameter :: SALMON_ERR_ARG = 1
  integer(c_int), parameter :: SALMON_ERR_HIP = 2
  integer(c_int), parameter :: SALMON_ERR_STATE = 3
  integer(c_int), parameter :: SALMON_ERR_NOMEM = 4
  integer(c_int), parameter :: SALMON_ERR_UNSUPPORTED = 5
  integer(c_int), parameter :: SALMON_FP64 = 0
  integer(c_int), parameter :: SALMON_MIXED = 1
  integer(c_int), parameter :: SALMON_PATH_AUTO = 0
  integer(c_int), parameter :: SALMON_PATH_STREAMED = 1
  integer(c_int), parameter :: SALMON_PATH_FUSED = 2
  integer(c_int), parameter :: SALMON_PATH_SWEEP = 3
  integer(c_int), parameter :: SALMON_PC_NONE = 0
  integer(c_int), parameter :: SALMON_PC_KINETIC = 1
  integer(c_int), parameter :: SALMON_PC_LEARNED = 2
  integer(c_int), parameter :: SALMON_PROP_MIDPOINT = 0
  integer(c_int), parameter :: SALMON_PROP_ETRS = 1
  integer(c_int), parameter :: SALMON_XC_NONE = 0
  integer(c_int), parameter :: SALMON_XC_PZ = 1

  type, bind(C) :: salmon_grid_t
    integer(c_int) :: NLx, NLy, NLz
    real(c_double) :: aLx, aLy, aLz
    real(c_double) :: lapx(0:4), lapy(0:4), lapz(0:4)
    real(c_double) :: nabx(0:4), naby(0:4), nabz(0:4)
  end type

  interface
    integer(c_int) function salmon_gpu_init(grid, precision, ctx) bind(C, name="salmon_gpu_init")
      import :: c_int, c_ptr, salmon_grid_t
      type(salmon_grid_t), intent(in) :: grid
      integer(c_int), value :: precision
      type(c_ptr), intent(out) :: ctx
    end function

    integer(c_int) function salmon_gpu_set_potential(ctx, vloc) bind(C, name="salmon_gpu_set_potential")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: vloc(*)
    end function

    integer(c_int) function salmon_gpu_set_projectors(ctx, NI, Nps, Mps, Jxyz, Jxx, Jyy, Jzz, Nlma, a_tbl, uV, iuV) &
        bind(C, name="salmon_gpu_set_projectors")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: NI, Nps, Nlma
      integer(c_int), intent(in) :: Mps(*), Jxyz(*), Jxx(*), Jyy(*), Jzz(*), a_tbl(*), iuV(*)
      real(c_double), intent(in) :: uV(*)
    end function

    integer(c_int) function salmon_gpu_set_kpoints(ctx, NK, NB, kAc0, occ) bind(C, name="salmon_gpu_set_kpoints")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: NK, NB
      real(c_double), intent(in) :: kAc0(*), occ(*)
    end function

    integer(c_int) function salmon_gpu_put_orbitals(ctx, zu) bind(C, name="salmon_gpu_put_orbitals")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(in) :: zu(*)
    end function

    integer(c_int) function salmon_gpu_get_orbitals(ctx, zu) bind(C, name="salmon_gpu_get_orbitals")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(out) :: zu(*)
    end function

    integer(c_int) function salmon_gpu_hpsi(ctx, kAc, tpsi, htpsi, ttpsi) bind(C, name="salmon_gpu_hpsi")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: kAc(3)
      type(c_ptr), value :: tpsi, htpsi, ttpsi
    end function

    integer(c_int) function salmon_gpu_hamiltonian(ctx, kAc, dt) bind(C, name="salmon_gpu_hamiltonian")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: kAc(*)
      real(c_double), value :: dt
    end function

    integer(c_int) function salmon_gpu_current(ctx, kAc, jav) bind(C, name="salmon_gpu_current")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: kAc(*)
      real(c_double), intent(out) :: jav(3)
    end function

    integer(c_int) function salmon_gpu_density(ctx, rho) bind(C, name="salmon_gpu_density")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: rho(*)
    end function

    integer(c_int) function salmon_gpu_energy(ctx, kAc, ekin, enl, eloc) bind(C, name="salmon_gpu_energy")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: kAc(*)
      real(c_double), intent(out) :: ekin, enl, eloc
    end function

    integer(c_int) function salmon_gpu_norm(ctx, norm) bind(C, name="salmon_gpu_norm")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: norm
    end function

    integer(c_int) function salmon_gpu_gram_schmidt(ctx) bind(C, name="salmon_gpu_gram_schmidt")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function salmon_gpu_sp_energy(ctx, kAc, esp) bind(C, name="salmon_gpu_sp_energy")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: kAc(*)
      real(c_double), intent(out) :: esp(*)
    end function

    integer(c_int) function salmon_gpu_cg(ctx, kAc, iter_cg_max, precond, esp_var, iters) &
        bind(C, name="salmon_gpu_cg")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: kAc(*)
      integer(c_int), value :: iter_cg_max, precond
      real(c_double), intent(out) :: esp_var(*)
      type(c_ptr), value :: iters
    end function

    integer(c_int) function salmon_gpu_subspace_diag(ctx, kAc, esp) bind(C, name="salmon_gpu_subspace_diag")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: kAc(*)
      real(c_double), intent(out) :: esp(*)
    end function

    integer(c_int) function salmon_gpu_ground_state(ctx, kAc, max_sweeps, iter_cg_max, precond, tol, esp, esp_var, &
        sweeps, history) bind(C, name="salmon_gpu_ground_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: kAc(*)
      integer(c_int), value :: max_sweeps, iter_cg_max, precond
      real(c_double), value :: tol
      real(c_double), intent(out) :: esp(*), esp_var(*)
      integer(c_int), intent(out) :: sweeps
      type(c_ptr), value :: history
    end function

    integer(c_int) function salmon_gpu_pcnet_create(seed, net) bind(C, name="salmon_gpu_pcnet_create")
      import :: c_int, c_ptr
      integer(c_int), value :: seed
      type(c_ptr), intent(out) :: net
    end function

    integer(c_int) function salmon_gpu_pcnet_set_weights(net, w) bind(C, name="salmon_gpu_pcnet_set_weights")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: net
      real(c_float), intent(in) :: w(*)
    end function

    integer(c_int) function salmon_gpu_pcnet_get_weights(net, w) bind(C, name="salmon_gpu_pcnet_get_weights")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: net
      real(c_float), intent(out) :: w(*)
    end function

    integer(c_int) function salmon_gpu_pcnet_finalize(net) bind(C, name="salmon_gpu_pcnet_finalize")
      import :: c_int, c_ptr
      type(c_ptr), value :: net
    end function

    integer(c_int) function salmon_gpu_pcnet_apply(net, ctx, g, d) bind(C, name="salmon_gpu_pcnet_apply")
      import :: c_int, c_ptr, c_double_complex
      type(c_ptr), value :: net, ctx
      complex(c_double_complex), intent(in) :: g(*)
      complex(c_double_complex), intent(out) :: d(*)
    end function

    integer(c_int) function salmon_gpu_set_propagator(ctx, propagator) bind(C, name="salmon_gpu_set_propagator")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: propagator
    end function

    integer(c_int) function salmon_gpu_set_scf(ctx, vpsl, xc) bind(C, name="salmon_gpu_set_scf")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: vpsl(*)
      integer(c_int), value :: xc
    end function

    integer(c_int) function salmon_gpu_hartree_xc(ctx, rho, xc, vh, vxc) bind(C, name="salmon_gpu_hartree_xc")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: rho(*)
      integer(c_int), value :: xc
      real(c_double), intent(out) :: vh(*), vxc(*)
    end function

    integer(c_int) function salmon_gpu_set_graphs(ctx, enable) bind(C, name="salmon_gpu_set_graphs")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: enable
    end function

    integer(c_int) function salmon_gpu_set_pcnet(ctx, net) bind(C, name="salmon_gpu_set_pcnet")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, net
    end function

    integer(c_int) function salmon_gpu_set_pc_handover(ctx, max_learned_sweeps, stall_ratio) &
        bind(C, name="salmon_gpu_set_pc_handover")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: max_learned_sweeps
      real(c_double), value :: stall_ratio
    end function

    integer(c_int) function salmon_gpu_pc_record(ctx, max_samples, stride) bind(C, name="salmon_gpu_pc_record")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: max_samples, stride
    end function

    integer(c_int) function salmon_gpu_pc_samples(ctx, stored, valid) bind(C, name="salmon_gpu_pc_samples")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: stored, valid
    end function

    integer(c_int) function salmon_gpu_pc_get_sample(ctx, i, g, x, band, theta, valid) &
        bind(C, name="salmon_gpu_pc_get_sample")
      import :: c_int, c_ptr, c_double, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: i
      complex(c_double_complex), intent(out) :: g(*), x(*)
      integer(c_int), intent(out) :: band, valid
      real(c_double), intent(out) :: theta
    end function

    integer(c_int) function salmon_gpu_pc_set_targets(ctx, kAc) bind(C, name="salmon_gpu_pc_set_targets")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: kAc(*)
    end function

    integer(c_int) function salmon_gpu_pcnet_gradient(net, ctx, nsamp, idx, loss, grad) &
        bind(C, name="salmon_gpu_pcnet_gradient")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: net, ctx
      integer(c_int), value :: nsamp
      integer(c_int), intent(in) :: idx(*)
      real(c_double), intent(out) :: loss, grad(*)
    end function

    integer(c_int) function salmon_gpu_pcnet_train(net, ctx, nsteps, batch, lr, seed, mean_loss) &
        bind(C, name="salmon_gpu_pcnet_train")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: net, ctx
      integer(c_int), value :: nsteps, batch, seed
      real(c_double), value :: lr
      real(c_double), intent(out) :: mean_loss
    end function

    integer(c_int) function salmon_gpu_rt_run(ctx, ac, nsteps, dt, reortho_every, jt, et, norm, rho) &
        bind(C, name="salmon_gpu_rt_run")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: ac(*)
      integer(c_int), value :: nsteps, reortho_every
      real(c_double), value :: dt
      type(c_ptr), value :: jt, et, norm, rho
    end function

    integer(c_int) function salmon_gpu_batch_create(precision, batch) bind(C, name="salmon_gpu_batch_create")
      import :: c_int, c_ptr
      integer(c_int), value :: precision
      type(c_ptr), intent(out) :: batch
    end function

    integer(c_int) function salmon_gpu_batch_add(batch, ctx) bind(C, name="salmon_gpu_batch_add")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch, ctx
    end function

    integer(c_int) function salmon_gpu_batch_set_kernel(batch, kernel) bind(C, name="salmon_gpu_batch_set_kernel")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
      integer(c_int), value :: kernel
    end function

    integer(c_int) function salmon_gpu_batch_kernel(batch) bind(C, name="salmon_gpu_batch_kernel")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
    end function

    integer(c_int) function salmon_gpu_batch_tune(batch) bind(C, name="salmon_gpu_batch_tune")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
    end function

    integer(c_int) function salmon_gpu_batch_zblock(batch, sys) bind(C, name="salmon_gpu_batch_zblock")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
      integer(c_int), value :: sys
    end function

    integer(c_int) function salmon_gpu_batch_rt_run(batch, ac, nsteps, dt, jt, et, norm) &
        bind(C, name="salmon_gpu_batch_rt_run")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: batch
      real(c_double), intent(in) :: ac(*)
      integer(c_int), value :: nsteps
      real(c_double), value :: dt
      type(c_ptr), value :: jt, et, norm
    end function

    ! ac(3, nsteps+1, nsys): one A(t) per macro-point
    integer(c_int) function salmon_gpu_batch_rt_run_ms(batch, ac, nsteps, dt, jt, et, norm) &
        bind(C, name="salmon_gpu_batch_rt_run_ms")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: batch
      real(c_double), intent(in) :: ac(*)
      integer(c_int), value :: nsteps
      real(c_double), value :: dt
      type(c_ptr), value :: jt, et, norm
    end function

    type(c_ptr) function salmon_gpu_batch_last_error(batch) bind(C, name="salmon_gpu_batch_last_error")
      import :: c_ptr
      type(c_ptr), value :: batch
    end function

    integer(c_int) function salmon_gpu_batch_finalize(batch) bind(C, name="salmon_gpu_batch_finalize")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
    end function

    integer(c_int) function salmon_gpu_set_path(ctx, path) bind(C, name="salmon_gpu_set_path")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: path
    end function

    integer(c_int) function salmon_gpu_path(ctx) bind(C, name="salmon_gpu_path")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function salmon_gpu_synchronize(ctx) bind(C, name="salmon_gpu_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    ! region timers; region ids are the LOG_* numbers of the reference's
    ! timer module (misc/timer.f90:21-54), e.g. LOG_DYNAMICS, LOG_HPSI
    integer(c_int) function salmon_gpu_timer_enable(ctx, on) bind(C, name="salmon_gpu_timer_enable")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
    end function

    integer(c_int) function salmon_gpu_timer_reset(ctx) bind(C, name="salmon_gpu_timer_reset")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function salmon_gpu_timer_get(ctx, id, seconds, calls) bind(C, name="salmon_gpu_timer_get")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: id
      real(c_double), intent(out) :: seconds
      integer(c_int), intent(out) :: calls
    end function

    type(c_ptr) function salmon_gpu_stream(ctx) bind(C, name="salmon_gpu_stream")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function

    type(c_ptr) function salmon_gpu_errstr(code) bind(C, name="salmon_gpu_errstr")
      import :: c_int, c_ptr
      integer(c_int), value :: code
    end function

    type(c_ptr) function salmon_gpu_last_error(ctx) bind(C, name="salmon_gpu_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function salmon_gpu_finalize(ctx) bind(C, name="salmon_gpu_finalize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
  end interface

contains

  ! The message behind a status code, as a Fortran string.
  function salmon_gpu_message(ctx, code) result(msg)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: code
    character(len=:), allocatable :: msg
    msg = c_string(salmon_gpu_errstr(code)) // ": " // c_string(salmon_gpu_last_error(ctx))
  end function

  function c_string(p) result(s)
    type(c_ptr), intent(in) :: p
    character(len=:), allocatable :: s
    character(kind=c_char), pointer :: chars(:)
    integer :: n
    interface
      integer(c_size_t) function strlen(cs) bind(C, name="strlen")
        import :: c_size_t, c_ptr
        type(c_ptr), value :: cs
      end function
    end interface
    if (.not. c_associated(p)) then
      s = ""
      return
    end if
    n = int(strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: s)
    s = transfer(chars(1:n), s)
  end function
end module salmon_gpu
