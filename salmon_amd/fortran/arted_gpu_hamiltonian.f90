! Drop-in for ARTED's  subroutine hamiltonian(zu, flag_current)
! (ARTED/RT/hamiltonian.f90:27-123): same external name, same arguments, same
! module state read (Global_Variables: grid, lapx..nabz, Vloc, projector
! tables, kAc, dt, occ).  A SALMON build selects this file in place of
! RT/hamiltonian.f90 the way it selects the OpenACC twin RT/acc/hamiltonian.f90
! (ARTED/RT/CMakeLists.txt:1-22).  The 4th-order Taylor step of every
! (k, band) orbital runs on the GPU through the C-ABI of include/salmon_gpu.h;
! a failed call stops through err_finalize, as the reference does on errors
! (ARTED/main.f90:21-33).
!
! The orbitals stay on the GPU between calls, as the OpenACC build keeps zu_t
! there ("!$acc enter data copyin(zu_t)", ARTED/control/control_sc.f90:149):
! the first call uploads zu, later calls on the same array find it resident,
! and the density, current and energy drop-ins (arted_gpu_observables.f90)
! read the device copy.  Host code that reads or writes zu_t itself marks the
! transfer as the OpenACC build does with "!$acc update self(zu_t)"
! (control_sc.f90:215,395) and "!$acc update device":
!   call arted_gpu_update_self(zu_t)    ! before a host read
!   call arted_gpu_update_device(zu_t)  ! after a host write
!   call arted_gpu_exit_data(zu_t)      ! end of the RT loop (control_sc.f90:414)
! A call with another orbital array (the multiscale loop's zu_m(:,:,:,imacro))
! first returns the resident array's orbitals to the host, then uploads the
! new one.
module arted_gpu_state
  use iso_c_binding
  use salmon_gpu
  implicit none
  type(c_ptr), save :: ctx = c_null_ptr
  logical, save :: ready = .false.
  ! the host array whose orbitals the device holds, and whether the device
  ! copy is newer than the host's
  type(c_ptr), save :: res_ptr = c_null_ptr
  logical, save :: host_stale = .false.
  integer(c_int), save, bind(C, name="arted_gpu_hamiltonian_calls") :: ncalls = 0
  ! whole-orbital-set transfers between host and device (uploads + downloads)
  integer(c_int), save, bind(C, name="arted_gpu_orbital_transfers") :: ntransfers = 0
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
         NK_s, NK_e, NBoccmax, kAc0, occ
    type(salmon_grid_t) :: g
    real(8), allocatable :: k0(:,:), oc(:,:)
    if (ready) return
    g%NLx = NLx; g%NLy = NLy; g%NLz = NLz
    g%aLx = aLx; g%aLy = aLy; g%aLz = aLz
    g%lapx = lapx(0:4); g%lapy = lapy(0:4); g%lapz = lapz(0:4)
    g%nabx = nabx(0:4); g%naby = naby(0:4); g%nabz = nabz(0:4)
    call arted_gpu_check(salmon_gpu_init(g, SALMON_FP64, ctx), 'init')
    call arted_gpu_tables()
    allocate(k0(NK_s:NK_e, 3), oc(NBoccmax, NK_s:NK_e))
    k0 = kAc0(NK_s:NK_e, :)
    oc = occ(1:NBoccmax, NK_s:NK_e)
    call arted_gpu_check(salmon_gpu_set_kpoints(ctx, NK_e - NK_s + 1, NBoccmax, k0, oc), 'set_kpoints')
    ready = .true.
  end subroutine

  ! The KB projector tables of the current atom positions (prep_ps_periodic
  ! rebuilds them every step of an Ehrenfest run, control_sc.f90:176-178).
  subroutine arted_gpu_tables()
    use Global_Variables, only: NI, Nps, Mps, Jxyz, Jxx, Jyy, Jzz, Nlma, a_tbl, uV, iuV
    call arted_gpu_check(salmon_gpu_set_projectors(ctx, NI, Nps, Mps, Jxyz, Jxx, Jyy, Jzz, Nlma, a_tbl, uV, iuV), &
         'set_projectors')
  end subroutine

  ! Makes the device hold the orbitals of zu ("!$acc enter data copyin").
  subroutine arted_gpu_enter(zu)
    use Global_Variables, only: NL, NBoccmax, NK_s, NK_e
    complex(8), intent(in), target :: zu(NL, NBoccmax, NK_s:NK_e)
    call arted_gpu_setup()
    if (c_associated(res_ptr, c_loc(zu))) return
    call arted_gpu_release()
    call arted_gpu_check(salmon_gpu_put_orbitals(ctx, zu), 'put_orbitals')
    ntransfers = ntransfers + 1
    res_ptr = c_loc(zu)
    host_stale = .false.
  end subroutine

  ! Returns newer device orbitals to the host array they belong to and forgets it.
  subroutine arted_gpu_release()
    use Global_Variables, only: NL, NBoccmax, NK_s, NK_e
    complex(8), pointer :: z(:)
    if (.not. c_associated(res_ptr)) return
    if (host_stale) then
      call c_f_pointer(res_ptr, z, [NL * NBoccmax * (NK_e - NK_s + 1)])
      call arted_gpu_check(salmon_gpu_get_orbitals(ctx, z), 'get_orbitals')
      ntransfers = ntransfers + 1
    end if
    res_ptr = c_null_ptr
    host_stale = .false.
  end subroutine

  ! "!$acc update self(zu)": the host copy of zu becomes current.
  subroutine arted_gpu_update_self(zu)
    use Global_Variables, only: NL, NBoccmax, NK_s, NK_e
    complex(8), intent(inout), target :: zu(NL, NBoccmax, NK_s:NK_e)
    if (.not. c_associated(res_ptr, c_loc(zu)) .or. .not. host_stale) return
    call arted_gpu_check(salmon_gpu_get_orbitals(ctx, zu), 'get_orbitals')
    ntransfers = ntransfers + 1
    host_stale = .false.
  end subroutine

  ! "!$acc update device(zu)": the host wrote zu; the device takes its copy.
  subroutine arted_gpu_update_device(zu)
    use Global_Variables, only: NL, NBoccmax, NK_s, NK_e
    complex(8), intent(in), target :: zu(NL, NBoccmax, NK_s:NK_e)
    if (.not. c_associated(res_ptr, c_loc(zu))) return
    call arted_gpu_check(salmon_gpu_put_orbitals(ctx, zu), 'put_orbitals')
    ntransfers = ntransfers + 1
    host_stale = .false.
  end subroutine

  ! "!$acc exit data copyout(zu)": the host copy becomes current and the
  ! device copy is no longer used.
  subroutine arted_gpu_exit_data(zu)
    use Global_Variables, only: NL, NBoccmax, NK_s, NK_e
    complex(8), intent(inout), target :: zu(NL, NBoccmax, NK_s:NK_e)
    if (.not. c_associated(res_ptr, c_loc(zu))) return
    call arted_gpu_update_self(zu)
    res_ptr = c_null_ptr
  end subroutine
end module arted_gpu_state

subroutine hamiltonian(zu, flag_current)
  use Global_Variables, only: NL, NBoccmax, NK_s, NK_e, kAc, dt, Vloc
  use salmon_global, only: use_ehrenfest_md, alocal_laser
  use arted_gpu_state
  implicit none
  complex(8), intent(inout), target :: zu(NL, NBoccmax, NK_s:NK_e)
  logical, intent(in) :: flag_current
  real(8), allocatable :: k(:,:)

  if (.false. .and. flag_current) return  ! current preprocessing is not part of this path
  if (alocal_laser == 'y') call err_finalize('GPU hamiltonian: alocal_laser is not built')
  call arted_gpu_setup()
  ncalls = ncalls + 1
  ! moved atoms: the tables prep_ps_periodic rebuilt this step
  if (use_ehrenfest_md == 'y') call arted_gpu_tables()
  allocate(k(NK_s:NK_e, 3))
  k = kAc(NK_s:NK_e, :)
  ! the potential may change between steps (Hartree/XC update, dt_evolve.f90:205-207)
  call arted_gpu_check(salmon_gpu_set_potential(ctx, Vloc), 'set_potential')
  call arted_gpu_enter(zu)
  call arted_gpu_check(salmon_gpu_hamiltonian(ctx, k, dt), 'hamiltonian')
  host_stale = .true.
end subroutine hamiltonian
