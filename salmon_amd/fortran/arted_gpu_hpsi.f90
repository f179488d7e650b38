! Drop-in for ARTED's  module hpsi  (ARTED/common/hpsi.f90:28-208): the same
! module name and public procedures with the same arguments,
!   hpsi_omp_KB_GS(ik, tpsi, ttpsi, htpsi)     tpsi, ttpsi, htpsi (NL)
!   hpsi_omp_KB_RT(ik, tpsi, htpsi)            padded (0:PNLz-1,0:PNLy-1,0:PNLx-1)
!   hpsi_omp_KB_base(ik, tpsi, htpsi, ttpsi)   padded, ttpsi optional
! with H = stencil + local + KB applied on the GPU (salmon_gpu_hpsi) at the
! wave vector kAc(ik,:) and the current Vloc.  The padded forms are the
! reference's layout when built with ARTED_STENCIL_PADDING (PNLy = NLy+1,
! opt_variables.f90:105-118); only the NL grid points of a padded array are
! read or written.  A SALMON build selects this file in place of
! common/hpsi.f90.  Errors stop through err_finalize.
module hpsi
  implicit none
contains
  subroutine hpsi_omp_KB_GS(ik, tpsi, ttpsi, htpsi)
    use Global_Variables, only: NL
    use iso_c_binding
    implicit none
    integer, intent(in) :: ik
    complex(8), intent(in), target :: tpsi(NL)
    complex(8), intent(out), target :: ttpsi(NL), htpsi(NL)
    call gpu_hpsi(ik, tpsi, htpsi, ttpsi, .true.)
  end subroutine

  subroutine hpsi_omp_KB_RT(ik, tpsi, htpsi)
    use opt_variables, only: PNLx, PNLy, PNLz
    implicit none
    integer, intent(in) :: ik
    complex(8), intent(in) :: tpsi(0:PNLz-1, 0:PNLy-1, 0:PNLx-1)
    complex(8), intent(out) :: htpsi(0:PNLz-1, 0:PNLy-1, 0:PNLx-1)
    call hpsi_omp_KB_base(ik, tpsi, htpsi)
  end subroutine

  subroutine hpsi_omp_KB_base(ik, tpsi, htpsi, ttpsi)
    use Global_Variables, only: NL, NLx, NLy, NLz
    use opt_variables, only: PNLx, PNLy, PNLz
    implicit none
    integer, intent(in) :: ik
    complex(8), intent(in) :: tpsi(0:PNLz-1, 0:PNLy-1, 0:PNLx-1)
    complex(8), intent(out) :: htpsi(0:PNLz-1, 0:PNLy-1, 0:PNLx-1)
    complex(8), intent(out), optional :: ttpsi(0:PNLz-1, 0:PNLy-1, 0:PNLx-1)
    complex(8), allocatable, target :: a(:,:,:), h(:,:,:), t(:,:,:)
    allocate(a(0:NLz-1, 0:NLy-1, 0:NLx-1), h(0:NLz-1, 0:NLy-1, 0:NLx-1), t(0:NLz-1, 0:NLy-1, 0:NLx-1))
    a = tpsi(0:NLz-1, 0:NLy-1, 0:NLx-1)
    call gpu_hpsi(ik, a, h, t, present(ttpsi))
    htpsi(0:NLz-1, 0:NLy-1, 0:NLx-1) = h
    if (present(ttpsi)) ttpsi(0:NLz-1, 0:NLy-1, 0:NLx-1) = t
  end subroutine

  ! H psi on the GPU (the RT drop-in's context holds the grid and tables).
  ! The reference calls hpsi_omp_KB_GS from inside OpenMP loops over k points
  ! (GS/sp_energy.f90:39-49); the context, its stream and staging buffers
  ! serve one caller at a time, so the calls are serialised.
  subroutine gpu_hpsi(ik, psi, hp, tt, with_tt)
    use Global_Variables, only: NL, kAc, Vloc
    use salmon_global, only: alocal_laser
    use iso_c_binding
    use salmon_gpu
    use arted_gpu_state
    implicit none
    integer, intent(in) :: ik
    complex(8), intent(in), target :: psi(NL)
    complex(8), intent(out), target :: hp(NL), tt(NL)
    logical, intent(in) :: with_tt
    real(8) :: k(3)
    type(c_ptr) :: ptt
    if (alocal_laser == 'y') call err_finalize('GPU hpsi: alocal_laser is not built')
    k = kAc(ik, :)
    ptt = c_null_ptr
    if (with_tt) ptt = c_loc(tt)
!$omp critical (arted_gpu_context)
    call arted_gpu_setup()
    call arted_gpu_check(salmon_gpu_set_potential(ctx, Vloc), 'set_potential')
    call arted_gpu_check(salmon_gpu_hpsi(ctx, k, c_loc(psi), c_loc(hp), ptt), 'hpsi')
!$omp end critical (arted_gpu_context)
  end subroutine
end module hpsi
