! Harness around the reference's own ARTED Fortran -- TEST INFRASTRUCTURE ONLY.
!
! Built by oracle/ref/Makefile together with the reference's source files
! (compiled where they lie in the SALMON tree) into oracle/_ref/libsalmon_ref.so.
! It fills the reference's module globals from the seeded synthetic system and
! calls the reference routines unchanged:
!   fd_coef               ARTED/preparation/fd_coef.f90:20
!   opt_vars_initialize_* ARTED/modules/opt_variables.f90:84-179
!   update_projector      ARTED/common/projector.f90:29
!   hpsi_omp_KB_GS        ARTED/common/hpsi.f90:33
!   hamiltonian           ARTED/RT/hamiltonian.f90:27
!   current_RT            ARTED/RT/current.f90:65
!   psi_rho_RT            ARTED/common/psi_rho.f90:40
!   Total_Energy_omp      ARTED/common/total_energy.f90:17
!   Gram_Schmidt          ARTED/GS/Gram_Schmidt.f90:20
!   sp_energy_omp         ARTED/GS/sp_energy.f90:20
!   CG_omp                ARTED/GS/CG.f90:20
! What the harness sets itself, restating the reference set-up it replaces:
!   Lx/Ly/Lz ............ ARTED/preparation/init.f90:28-37
!   ik_table/ib_table ... ARTED/preparation/init.f90:101-109
!   communicators ....... ARTED/control/initialization.f90:340-342 (no MPI)
!   the RT step order ... ARTED/control/control_sc.f90:190-222 with
!                         ARTED/RT/dt_evolve.f90:43-51 (middlepoint) and a
!                         frozen potential (no_update_func='y').
!   moved atoms ......... ARTED/preparation/prep_ps.f90 ('update_all'), the
!                         state control_sc.f90:152-178 leaves in an
!                         Ehrenfest step (ref_move_atoms).
module ref_harness
  use iso_c_binding
#ifdef ARTED_GPU_DROPIN
  ! host reads and writes of zu_t mark their transfers, as the reference's
  ! OpenACC build does with "!$acc update self/device(zu_t)"
  use arted_gpu_state, only: arted_gpu_update_self, arted_gpu_update_device
#endif
  use Global_Variables
  use opt_variables
  use timer, only: timer_initialize
  use projector, only: update_projector
  use hpsi, only: hpsi_omp_KB_GS
  use omp_lib, only: omp_get_max_threads
  use salmon_parallel, only: setup_parallel, nproc_group_global, nproc_group_tdks
  use salmon_communication, only: comm_create_group
  implicit none
  real(8), allocatable :: rho_sum(:)
contains

  subroutine ref_setup(nlx_, nly_, nlz_, alx_, aly_, alz_, nk_, nb_, kac0_, occ_, vloc_, &
       ni_, nps_, mps_, jxyz_, jxx_, jyy_, jzz_, nlma_, atbl_, uv_, iuv_, rion_, rc_, dt_, status) &
       bind(C, name="ref_setup")
    integer(c_int), value :: nlx_, nly_, nlz_, nk_, nb_, ni_, nps_, nlma_
    real(c_double), value :: alx_, aly_, alz_, rc_, dt_
    real(c_double), intent(in) :: kac0_(nk_, 3), occ_(nb_, nk_), vloc_(nlx_*nly_*nlz_)
    integer(c_int), intent(in) :: mps_(ni_), jxyz_(nps_, ni_), jxx_(nps_, ni_), jyy_(nps_, ni_), jzz_(nps_, ni_)
    integer(c_int), intent(in) :: atbl_(nlma_), iuv_(nlma_)
    real(c_double), intent(in) :: uv_(nps_, nlma_), rion_(3, ni_)
    integer(c_int), intent(out) :: status
    integer :: n, ix, iy, iz, a, j, i, ik, ib, nth

    status = 0
    ! communicators without MPI, as main.f90 / initialization.f90:340-342
    call setup_parallel
    nproc_group_tdks = comm_create_group(nproc_group_global, 0, 0)
    NLx = nlx_; NLy = nly_; NLz = nlz_; NL = NLx*NLy*NLz
    aLx = alx_; aLy = aly_; aLz = alz_; aLxyz = aLx*aLy*aLz
    Hx = aLx/NLx; Hy = aLy/NLy; Hz = aLz/NLz; Hxyz = Hx*Hy*Hz
    bLx = 2*Pi/aLx; bLy = 2*Pi/aLy; bLz = 2*Pi/aLz
    allocate(lap(-Nd:Nd), nab(-Nd:Nd), lapx(-Nd:Nd), lapy(-Nd:Nd), lapz(-Nd:Nd))
    allocate(nabx(-Nd:Nd), naby(-Nd:Nd), nabz(-Nd:Nd))
    call fd_coef

    allocate(Lx(NL), Ly(NL), Lz(NL))
    n = 0
    do ix = 0, NLx-1
    do iy = 0, NLy-1
    do iz = 0, NLz-1
      n = n + 1
      Lx(n) = ix; Ly(n) = iy; Lz(n) = iz
    end do
    end do
    end do

    NK = nk_; NB = nb_; NBoccmax = nb_; NK_s = 1; NK_e = NK; NKB = NK*NB
    NKx = 0; NKy = 0; NKz = 0; NKxyz = NK
    allocate(kAc(NK,3), kAc0(NK,3), kAc_new(NK,3), wk(NK), occ(NB,NK))
    kAc0 = kac0_; kAc = kac0_; kAc_new = kac0_; wk = 1.d0; occ = occ_
    allocate(ik_table(NKB), ib_table(NKB))
    i = 0
    do ik = NK_s, NK_e
      do ib = 1, NBoccmax
        i = i + 1
        ik_table(i) = ik; ib_table(i) = ib
      end do
    end do

    allocate(Vloc(NL), Vpsl(NL), Vh(NL), Vexc(NL), Eexc(NL))
    Vloc = vloc_; Vpsl = vloc_; Vh = 0.d0; Vexc = 0.d0; Eexc = 0.d0

    NI = ni_; natom = ni_; Nps = nps_; Nlma = nlma_
    allocate(kion(NI), rion(3,NI), Rps(1))
    kion = 1; rion = rion_; Rps(1) = rc_
    allocate(Mps(NI), Jxyz(Nps,NI), Jxx(Nps,NI), Jyy(Nps,NI), Jzz(Nps,NI))
    Mps = mps_; Jxyz = jxyz_; Jxx = jxx_; Jyy = jyy_; Jzz = jzz_
    allocate(a_tbl(Nlma), uV(Nps,Nlma), iuV(Nlma))
    a_tbl = atbl_; uV = uv_; iuV = iuv_
    allocate(zproj(Nps,Nlma,NK_s:NK_e), ekr_omp(Nps,NI,NK_s:NK_e))

    dt = dt_
    use_ehrenfest_md = 'n'
    functional = 'PZ'
    Sym = 1
    alocal_laser = 'n'
    skip_gsortho = 'n'
    omp_loop = 'k'
    NUMBER_THREADS = omp_get_max_threads()
    nth = NUMBER_THREADS

    call timer_initialize
    call opt_vars_initialize_p1
    call opt_vars_initialize_p2
    ! opt_vars_initialize_p2 rebuilds the padded projector index from the atom
    ! positions (init_for_padding); it must agree with the tables passed in.
    do a = 1, NI
      do j = 1, Mps(a)
        i = Jxyz(j,a)
        if (zKxyz(j,a) /= Lx(i)*PNLy*PNLz + Ly(i)*PNLz + Lz(i)) status = 1
      end do
    end do

    allocate(zu_t(NL,NB,NK), zu_GS(NL,NB,NK))
    zu_t = (0.d0, 0.d0); zu_GS = (0.d0, 0.d0)
    allocate(rho(NL), rho_l(NL), rho_sum(NL), esp(NB,NK), esp_var(NB,NK))
    allocate(tpsi_omp(NL,0:nth-1), ttpsi_omp(NL,0:nth-1), htpsi_omp(NL,0:nth-1))
    allocate(xk_omp(NL,0:nth-1), hxk_omp(NL,0:nth-1), gk_omp(NL,0:nth-1), pk_omp(NL,0:nth-1))
    allocate(pko_omp(NL,0:nth-1), txk_omp(NL,0:nth-1))
    ! G-space parts of Total_Energy_omp are outside this path: empty G range
    NG = 0; NG_s = 1; NG_e = 0; nGzero = 0
    allocate(Gx(0), Gy(0), Gz(0), rhoe_G(0), rhoion_G(0), dVloc_G(0,1))
  end subroutine

  ! Moved atoms in an Ehrenfest run: the module state prep_ps_periodic
  ! ('update_all') leaves for the next step (ARTED/preparation/prep_ps.f90:
  ! 54-66 projector index freed, 95-116 Nps-sized arrays reallocated when Nps
  ! changes, 118-144 Jxyz/Jxx/Jyy/Jzz and the padded index zKxyz, 253-256
  ! init_projector), with the sphere tables, uV and the local potential of
  ! the new positions passed in (calc_jxyz / calc_uv / calc_vpsl are outside
  ! this path).  Same atoms and projectors (NI, Nlma, a_tbl, iuV unchanged).
  subroutine ref_move_atoms(nps_, mps_, jxyz_, jxx_, jyy_, jzz_, uv_, rion_, vloc_) bind(C, name="ref_move_atoms")
    integer(c_int), value :: nps_
    integer(c_int), intent(in) :: mps_(NI), jxyz_(nps_, NI), jxx_(nps_, NI), jyy_(nps_, NI), jzz_(nps_, NI)
    real(c_double), intent(in) :: uv_(nps_, Nlma), rion_(3, NI), vloc_(NL)
    integer :: a, j, i

    use_ehrenfest_md = 'y'
    if (nps_ /= Nps) then
      deallocate(Jxyz, Jxx, Jyy, Jzz, zJxyz, zKxyz, ekr_omp, uV, zproj)
      Nps = nps_
      allocate(Jxyz(Nps,NI), Jxx(Nps,NI), Jyy(Nps,NI), Jzz(Nps,NI), zJxyz(Nps,NI), zKxyz(Nps,NI))
      allocate(ekr_omp(Nps,NI,NK_s:NK_e), uV(Nps,Nlma), zproj(Nps,Nlma,NK_s:NK_e))
    end if
    Mps = mps_; Jxyz = jxyz_; Jxx = jxx_; Jyy = jyy_; Jzz = jzz_; uV = uv_
    Rion = rion_; Vloc = vloc_; Vpsl = vloc_
    zJxyz = Jxyz - 1
    zKxyz = 0
    do a = 1, NI
      do j = 1, Mps(a)
        i = Jxyz(j,a)
        zKxyz(j,a) = Lx(i)*PNLy*PNLz + Ly(i)*PNLz + Lz(i)
      end do
    end do
    deallocate(nprojector, idx_proj, idx_lma, pseudo_start_idx)
    call init_projector(zKxyz)
  end subroutine

  subroutine ref_get_fd(out) bind(C, name="ref_get_fd")
    real(c_double), intent(out) :: out(0:4, 6)
    out(:,1) = lapx(0:4); out(:,2) = lapy(0:4); out(:,3) = lapz(0:4)
    out(:,4) = nabx(0:4); out(:,5) = naby(0:4); out(:,6) = nabz(0:4)
  end subroutine

  subroutine ref_put_orbitals(zu) bind(C, name="ref_put_orbitals")
    complex(c_double_complex), intent(in) :: zu(NL, NB, NK)
    zu_t = zu
#ifdef ARTED_GPU_DROPIN
    call arted_gpu_update_device(zu_t)
#endif
  end subroutine

  subroutine ref_get_orbitals(zu) bind(C, name="ref_get_orbitals")
    complex(c_double_complex), intent(out) :: zu(NL, NB, NK)
#ifdef ARTED_GPU_DROPIN
    call arted_gpu_update_self(zu_t)
#endif
    zu = zu_t
  end subroutine

  subroutine ref_hpsi(ik, kvec, tpsi_, htpsi_, ttpsi_) bind(C, name="ref_hpsi")
    integer(c_int), value :: ik
    real(c_double), intent(in) :: kvec(3)
    complex(c_double_complex), intent(in) :: tpsi_(NL)
    complex(c_double_complex), intent(out) :: htpsi_(NL), ttpsi_(NL)
    kAc(ik,:) = kvec
    call update_projector(kAc)
    call hpsi_omp_KB_GS(ik, tpsi_, ttpsi_, htpsi_)
  end subroutine

  subroutine ref_hamiltonian(kac_) bind(C, name="ref_hamiltonian")
    real(c_double), intent(in) :: kac_(NK, 3)
    interface
      subroutine hamiltonian(zu, flag_current)
        complex(8), intent(inout) :: zu(*)
        logical, intent(in) :: flag_current
      end subroutine
    end interface
    kAc = kac_
    call update_projector(kAc)
    call hamiltonian(zu_t, .true.)
  end subroutine

  subroutine ref_current(kac_, jout) bind(C, name="ref_current")
    real(c_double), intent(in) :: kac_(NK, 3)
    real(c_double), intent(out) :: jout(3)
    kAc = kac_
    call current_RT
    jout = jav
  end subroutine

  subroutine ref_psi_rho(rout) bind(C, name="ref_psi_rho")
    real(c_double), intent(out) :: rout(NL)
    interface
      subroutine psi_rho_RT(zu)
        complex(8), intent(in) :: zu(*)
      end subroutine
    end interface
    call psi_rho_RT(zu_t)
    rout = rho
  end subroutine

  subroutine ref_energy(kac_, ekin_, enl_) bind(C, name="ref_energy")
    real(c_double), intent(in) :: kac_(NK, 3)
    real(c_double), intent(out) :: ekin_, enl_
    interface
      subroutine Total_Energy_omp(Rion_update, GS_RT, ixy_m)
        logical, intent(in) :: Rion_update
        integer, intent(in) :: GS_RT
        integer, intent(in), optional :: ixy_m
      end subroutine
    end interface
    kAc = kac_
    call Total_Energy_omp(.false., calc_mode_rt)
    ekin_ = Ekin; enl_ = Enl
  end subroutine

  subroutine ref_gram_schmidt() bind(C, name="ref_gram_schmidt")
#ifdef ARTED_GPU_DROPIN
    call arted_gpu_update_self(zu_t)
#endif
    zu_GS = zu_t
    call Gram_Schmidt
    zu_t = zu_GS
#ifdef ARTED_GPU_DROPIN
    call arted_gpu_update_device(zu_t)
#endif
  end subroutine

  subroutine ref_sp_energy(kac_, esp_out) bind(C, name="ref_sp_energy")
    real(c_double), intent(in) :: kac_(NK, 3)
    real(c_double), intent(out) :: esp_out(NB, NK)
    kAc = kac_
#ifdef ARTED_GPU_DROPIN
    call arted_gpu_update_self(zu_t)
#endif
    zu_GS = zu_t
    call sp_energy_omp
    esp_out = esp
  end subroutine

  subroutine ref_cg(kac_, ncg, esp_var_out) bind(C, name="ref_cg")
    real(c_double), intent(in) :: kac_(NK, 3)
    integer(c_int), value :: ncg
    real(c_double), intent(out) :: esp_var_out(NB, NK)
    kAc = kac_
#ifdef ARTED_GPU_DROPIN
    call arted_gpu_update_self(zu_t)
#endif
    zu_GS = zu_t
    call CG_omp(ncg)
    zu_t = zu_GS
#ifdef ARTED_GPU_DROPIN
    call arted_gpu_update_device(zu_t)
#endif
    esp_var_out = esp_var
  end subroutine

  ! One RT run of nsteps from the orbitals in zu_t.  ac(3,0:nsteps) holds
  ! A(t_n).  Per step: Taylor-4 at the midpoint k+A, density, then current
  ! and (every energy_every steps) energy at k+A(t+dt).
  subroutine ref_rt_run(ac, nsteps, jt, et, energy_every) bind(C, name="ref_rt_run")
    integer(c_int), value :: nsteps, energy_every
    real(c_double), intent(in) :: ac(3, 0:nsteps)
    real(c_double), intent(out) :: jt(3, nsteps), et(nsteps)
    integer :: it, d
    interface
      subroutine hamiltonian(zu, flag_current)
        complex(8), intent(inout) :: zu(*)
        logical, intent(in) :: flag_current
      end subroutine
      subroutine psi_rho_RT(zu)
        complex(8), intent(in) :: zu(*)
      end subroutine
      subroutine Total_Energy_omp(Rion_update, GS_RT, ixy_m)
        logical, intent(in) :: Rion_update
        integer, intent(in) :: GS_RT
        integer, intent(in), optional :: ixy_m
      end subroutine
    end interface
    do it = 0, nsteps-1
      do d = 1, 3
        kAc(:,d) = kAc0(:,d) + 0.5d0*(ac(d,it) + ac(d,it+1))
      end do
      call update_projector(kAc)
      call hamiltonian(zu_t, .true.)
      call psi_rho_RT(zu_t)
      do d = 1, 3
        kAc(:,d) = kAc0(:,d) + ac(d,it+1)
      end do
      call current_RT
      jt(:, it+1) = jav
      et(it+1) = 0.d0
      if (energy_every <= 1 .or. mod(it+1, max(energy_every,1)) == 0) then
        call Total_Energy_omp(.false., calc_mode_rt)
        et(it+1) = Ekin + Enl + sum(Vloc*rho)*Hxyz
      end if
    end do
  end subroutine

  integer(c_int) function ref_num_threads() bind(C, name="ref_num_threads")
    ref_num_threads = omp_get_max_threads()
  end function
end module ref_harness

! The reference's err_finalize (ARTED/main.f90:21-33) lives in the file of the
! full program driver; this harness stops the same way without MPI.
subroutine err_finalize(err_message)
  implicit none
  character(*), intent(in) :: err_message
  write(*,*) err_message
  stop
end subroutine err_finalize
