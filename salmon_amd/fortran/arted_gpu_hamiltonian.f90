! Drop-in for ARTED's  subroutine hamiltonian(zu, flag_current)
! (ARTED/RT/hamiltonian.f90:27-123): same external name, same arguments, same
! module state read (Global_Variables: grid, lapx..nabz, Vloc, projector
! tables, kAc, dt, occ).  A SALMON build selects this file in place of
! RT/hamiltonian.f90 the way it selects the OpenACC twin RT/acc/hamiltonian.f90
! (ARTED/RT/CMakeLists.txt:1-22).  The 4th-order Taylor step of every
! (k, band) orbital runs on the GPU through the C-ABI of include/salmon_gpu.h;
! a failed call stops through err_finalize, as the reference does on errors
! (ARTED/main.f90:21-33).
module arted_gpu_state
  use iso_c_binding
  use salmon_gpu
  implicit none
  type(c_ptr), save :: ctx = c_null_ptr
  logical, save :: ready = .false.
  integer(c_int), save, bind(C, name="arted_gpu_hamiltonian_calls") :: ncalls = 0
contains
  subroutine arted_gpu_check(st, what)
    integer(c_int), intent(in) :: st
    character(*), intent(in) :: what
    if (st /= SALMON_OK) call err_finalize('salmon_gpu '//what//': '//salmon_gpu_message(ctx, st))
  end subroutine

  ! Device copies of the module state that does not change between steps
  ! (grid, FD weights, projector tables, k points); set up on the first call.
  subroutine arted_gpu_setup()
    use Global_Variables, only: NLx, NLy, NLz, aLx, aLy, aLz, lapx, lapy, lapz, nabx, naby, nabz, &
         NI, Nps, Mps, Jxyz, Jxx, Jyy, Jzz, Nlma, a_tbl, uV, iuV, NK_s, NK_e, NBoccmax, kAc0, occ
    type(salmon_grid_t) :: g
    real(8), allocatable :: k0(:,:), oc(:,:)
    if (ready) return
    g%NLx = NLx; g%NLy = NLy; g%NLz = NLz
    g%aLx = aLx; g%aLy = aLy; g%aLz = aLz
    g%lapx = lapx(0:4); g%lapy = lapy(0:4); g%lapz = lapz(0:4)
    g%nabx = nabx(0:4); g%naby = naby(0:4); g%nabz = nabz(0:4)
    call arted_gpu_check(salmon_gpu_init(g, SALMON_FP64, ctx), 'init')
    call arted_gpu_check(salmon_gpu_set_projectors(ctx, NI, Nps, Mps, Jxyz, Jxx, Jyy, Jzz, Nlma, a_tbl, uV, iuV), &
         'set_projectors')
    allocate(k0(NK_s:NK_e, 3), oc(NBoccmax, NK_s:NK_e))
    k0 = kAc0(NK_s:NK_e, :)
    oc = occ(1:NBoccmax, NK_s:NK_e)
    call arted_gpu_check(salmon_gpu_set_kpoints(ctx, NK_e - NK_s + 1, NBoccmax, k0, oc), 'set_kpoints')
    ready = .true.
  end subroutine
end module arted_gpu_state

subroutine hamiltonian(zu, flag_current)
  use Global_Variables, only: NL, NBoccmax, NK_s, NK_e, kAc, dt, Vloc
  use arted_gpu_state
  implicit none
  complex(8), intent(inout) :: zu(NL, NBoccmax, NK_s:NK_e)
  logical, intent(in) :: flag_current
  real(8), allocatable :: k(:,:)

  if (.false. .and. flag_current) return  ! current preprocessing is not part of this path
  call arted_gpu_setup()
  ncalls = ncalls + 1
  allocate(k(NK_s:NK_e, 3))
  k = kAc(NK_s:NK_e, :)
  ! the potential may change between steps (Hartree/XC update, dt_evolve.f90:205-207)
  call arted_gpu_check(salmon_gpu_set_potential(ctx, Vloc), 'set_potential')
  call arted_gpu_check(salmon_gpu_put_orbitals(ctx, zu), 'put_orbitals')
  call arted_gpu_check(salmon_gpu_hamiltonian(ctx, k, dt), 'hamiltonian')
  call arted_gpu_check(salmon_gpu_get_orbitals(ctx, zu), 'get_orbitals')
end subroutine hamiltonian
