! Drop-ins for the observables of ARTED's RT step, reading the orbitals the
! GPU holds (see arted_gpu_hamiltonian.f90) instead of the host's zu_t:
!
!   subroutine current_acc_impl(zutmp, jx, jy, jz)
!       the reference's own GPU swap point of the current: current()
!       (ARTED/RT/current.f90:78-104) calls it for mode 'RT' when built with
!       _OPENACC, and RT/acc/current.f90 provides it (selected with
!       acc/hamiltonian.f90, ARTED/RT/CMakeLists.txt:8-11).  A SALMON build
!       compiles RT/current.f90 with -D_OPENACC and this file in place of
!       RT/acc/current.f90; current_RT, current_RT_MS and the summation over
!       processes stay the reference's.
!   subroutine psi_rho_RT(zu), psi_rho_GS
!       in place of ARTED/common/psi_rho.f90 (:34-46): rho = sum occ |zu|^2
!       (Sym = 1; the symmetrised Sym = 4/8 forms are outside this path and stop
!       through err_finalize), summed over processes as psi_rho_impl does
!       (:86-88).
!   subroutine Total_Energy_omp(Rion_update, GS_RT, ixy_m)
!       in place of ARTED/common/total_energy.f90: the wavefunction terms Ekin
!       and Enl (:113-203) on the GPU; the ion-ion Ewald sum and the G-space
!       Hartree and local terms (:71-103, :124-140) and the totals (:205-231) on
!       the host in the reference's formulas.
!
! Errors stop through err_finalize (ARTED/main.f90:21-33).

subroutine current_acc_impl(zutmp, jxs, jys, jzs)
  use Global_Variables, only: NL, NBoccmax, NK_s, NK_e, kAc
  use arted_gpu_state
  implicit none
  complex(8), intent(in), target :: zutmp(0:NL-1, NBoccmax, NK_s:NK_e)
  real(8), intent(out) :: jxs, jys, jzs
  real(8), allocatable :: k(:,:)
  real(8) :: j(3)
  call arted_gpu_enter(zutmp)
  allocate(k(NK_s:NK_e, 3))
  k = kAc(NK_s:NK_e, :)
  call arted_gpu_check(salmon_gpu_current(ctx, k, j), 'current')
  jxs = j(1); jys = j(2); jzs = j(3)
end subroutine current_acc_impl

subroutine psi_rho_RT(zu)
  use Global_Variables, only: NL, NBoccmax, NK_s, NK_e, Sym, rho, rho_l
  use salmon_parallel, only: nproc_group_tdks
  use salmon_communication, only: comm_summation
  use arted_gpu_state
  implicit none
  complex(8), intent(in), target :: zu(NL, NBoccmax, NK_s:NK_e)
  if (Sym /= 1) call err_finalize('GPU psi_rho_RT: Sym = 1 only (symmetrised densities are not built)')
  call arted_gpu_enter(zu)
  call arted_gpu_check(salmon_gpu_density(ctx, rho_l), 'density')
  call comm_summation(rho_l, rho, NL, nproc_group_tdks)
end subroutine psi_rho_RT

subroutine psi_rho_GS
  use Global_Variables, only: NL, Sym, rho, rho_l, zu_GS
  use salmon_parallel, only: nproc_group_tdks
  use salmon_communication, only: comm_summation
  use arted_gpu_gs_state
  implicit none
  if (Sym /= 1) call err_finalize('GPU psi_rho_GS: Sym = 1 only (symmetrised densities are not built)')
  call arted_gpu_gs_setup()
  call arted_gpu_gs_check(salmon_gpu_put_orbitals(gctx, zu_GS), 'put_orbitals')
  call arted_gpu_gs_check(salmon_gpu_density(gctx, rho_l), 'density')
  call comm_summation(rho_l, rho, NL, nproc_group_tdks)
end subroutine psi_rho_GS

subroutine Total_Energy_omp(Rion_update, GS_RT, ixy_m)
  use Global_Variables, only: NI, NEwald, aEwald, Rion, Zps, Kion, aLx, aLy, aLz, aLxyz, Pi, zI, NG_s, NG_e, &
       nGzero, Gx, Gy, Gz, rhoion_G, rhoe_G, dVloc_G, NK_s, NK_e, kAc, zu_t, zu_GS, calc_mode_gs, calc_mode_rt, &
       Ekin, Eh, Enl, Eloc, Exc, Eion, Eall, Eexc, Hxyz
  use salmon_math, only: erfc_salmon
  use salmon_global, only: alocal_laser
  use salmon_parallel, only: nproc_group_tdks
  use salmon_communication, only: comm_summation
  use arted_gpu_state
  use arted_gpu_gs_state
  implicit none
  logical, intent(in) :: Rion_update
  integer, intent(in) :: GS_RT
  integer, intent(in), optional :: ixy_m
  integer :: ia, ib, ix, iy, iz, n
  real(8) :: rab(3), rab2, G2, Gd, ek, en, el
  real(8) :: Eion_tmp1, Eion_tmp2, Eion_l, Eh_l, Eloc_l1, sum_tmp(5), sum_result(5)
  complex(8) :: Eloc_l2
  real(8), allocatable :: k(:,:)

  if (alocal_laser == 'y') call err_finalize('GPU Total_Energy_omp: alocal_laser is not built')
  if (present(ixy_m)) call err_finalize('GPU Total_Energy_omp: the multiscale orbitals zu_m run through '// &
       'salmon_gpu_batch_rt_run_ms')

  ! ion-ion: Ewald sum (total_energy.f90:71-103)
  if (Rion_update) then
    Eion_tmp1 = 0.d0
    Eion_l = 0.d0
    do ia = 1, NI
      do ix = -NEwald, NEwald
      do iy = -NEwald, NEwald
      do iz = -NEwald, NEwald
        do ib = 1, NI
          if (ix**2 + iy**2 + iz**2 == 0 .and. ia == ib) cycle
          rab(1) = Rion(1, ia) - ix*aLx - Rion(1, ib)
          rab(2) = Rion(2, ia) - iy*aLy - Rion(2, ib)
          rab(3) = Rion(3, ia) - iz*aLz - Rion(3, ib)
          rab2 = sum(rab(:)**2)
          Eion_tmp1 = Eion_tmp1 + 0.5d0*Zps(Kion(ia))*Zps(Kion(ib))*erfc_salmon(sqrt(aEwald*rab2))/sqrt(rab2)
        end do
      end do
      end do
      end do
    end do
    do n = NG_s, NG_e
      if (n == nGzero) cycle
      G2 = Gx(n)**2 + Gy(n)**2 + Gz(n)**2
      Eion_l = Eion_l + aLxyz*(4*Pi/G2)*(abs(rhoion_G(n))**2*exp(-G2/(4*aEwald))*0.5d0)
    end do
    Eion_tmp1 = Eion_tmp1 - Pi*sum(Zps(Kion(:)))**2/(2*aEwald*aLxyz) - sqrt(aEwald/Pi)*sum(Zps(Kion(:))**2)
  end if

  ! G-space local and Hartree terms (total_energy.f90:124-140)
  Eloc_l2 = 0.d0
  do ia = 1, NI
    do n = NG_s, NG_e
      Gd = Gx(n)*Rion(1, ia) + Gy(n)*Rion(2, ia) + Gz(n)*Rion(3, ia)
      Eloc_l2 = Eloc_l2 + conjg(rhoe_G(n))*dVloc_G(n, Kion(ia))*exp(-zI*Gd)
    end do
  end do
  Eh_l = 0.d0
  Eloc_l1 = 0.d0
  do n = NG_s, NG_e
    if (n == nGzero) cycle
    G2 = Gx(n)**2 + Gy(n)**2 + Gz(n)**2
    Eh_l = Eh_l + aLxyz*(4*Pi/G2)*(abs(rhoe_G(n))**2*0.5d0)
    Eloc_l1 = Eloc_l1 + aLxyz*(4*Pi/G2)*(-rhoe_G(n)*conjg(rhoion_G(n)))
  end do

  ! wavefunction terms on the GPU at kAc (total_energy.f90:164-203)
  allocate(k(NK_s:NK_e, 3))
  k = kAc(NK_s:NK_e, :)
  select case (GS_RT)
  case (calc_mode_rt)
    call arted_gpu_enter(zu_t)
    call arted_gpu_check(salmon_gpu_energy(ctx, k, ek, en, el), 'energy')
  case (calc_mode_gs)
    call arted_gpu_gs_setup()
    call arted_gpu_gs_check(salmon_gpu_put_orbitals(gctx, zu_GS), 'put_orbitals')
    call arted_gpu_gs_check(salmon_gpu_energy(gctx, k, ek, en, el), 'energy')
  case default
    call err_finalize('total_energy_omp: GS_RT flag')
  end select

  ! totals over the processes (total_energy.f90:205-231)
  if (Rion_update) then
    call comm_summation(Eion_l, Eion_tmp2, nproc_group_tdks)
    Eion = Eion_tmp1 + Eion_tmp2
  end if
  sum_tmp(1) = ek
  sum_tmp(2) = Eh_l
  sum_tmp(3) = en
  sum_tmp(4) = Eloc_l1
  sum_tmp(5) = real(Eloc_l2, 8)
  call comm_summation(sum_tmp, sum_result, 5, nproc_group_tdks)
  Ekin = sum_result(1)
  Eh = sum_result(2)
  Enl = sum_result(3)
  Eloc = sum_result(4) + sum_result(5)
  Exc = sum(Eexc)*Hxyz
  Eall = Ekin + Eloc + Enl + Exc + Eh + Eion
end subroutine Total_Energy_omp
