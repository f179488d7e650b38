! Drop-ins for ARTED's ground-state routines with the reference's external
! names and arguments:
!   subroutine CG_omp(iter_cg_max)   ARTED/GS/CG.f90:20-32 (-> CG_ompk :34-139)
!   subroutine Gram_Schmidt          ARTED/GS/Gram_Schmidt.f90:20-31
! reading and writing the same module state (Global_Variables: zu_GS, esp_var,
! kAc, Vloc, occ, the grid and projector tables).  A SALMON build selects this
! file in place of GS/CG.f90 and GS/Gram_Schmidt.f90, as it selects the GPU
! hamiltonian in place of RT/hamiltonian.f90.  The band-by-band CG and the
! Gram-Schmidt run on the GPU through the C-ABI of include/salmon_gpu.h (the
! reference's method: no preconditioner); errors stop through err_finalize
! (ARTED/main.f90:21-33).
module arted_gpu_gs_state
  use iso_c_binding
  use salmon_gpu
  implicit none
  type(c_ptr), save :: gctx = c_null_ptr
  logical, save :: gready = .false.
  integer(c_int), save, bind(C, name="arted_gpu_gs_calls") :: ngs_calls = 0
contains
  subroutine arted_gpu_gs_check(st, what)
    integer(c_int), intent(in) :: st
    character(*), intent(in) :: what
    if (st /= SALMON_OK) call err_finalize('salmon_gpu '//what//': '//salmon_gpu_message(gctx, st))
  end subroutine

  ! A context with the ground state's NB bands (the RT drop-in's has NBoccmax).
  subroutine arted_gpu_gs_setup()
    use Global_Variables, only: NLx, NLy, NLz, aLx, aLy, aLz, lapx, lapy, lapz, nabx, naby, nabz, &
         NI, Nps, Mps, Jxyz, Jxx, Jyy, Jzz, Nlma, a_tbl, uV, iuV, NK_s, NK_e, NB, kAc0, occ
    type(salmon_grid_t) :: g
    real(8), allocatable :: k0(:,:), oc(:,:)
    if (gready) return
    g%NLx = NLx; g%NLy = NLy; g%NLz = NLz
    g%aLx = aLx; g%aLy = aLy; g%aLz = aLz
    g%lapx = lapx(0:4); g%lapy = lapy(0:4); g%lapz = lapz(0:4)
    g%nabx = nabx(0:4); g%naby = naby(0:4); g%nabz = nabz(0:4)
    call arted_gpu_gs_check(salmon_gpu_init(g, SALMON_FP64, gctx), 'init')
    call arted_gpu_gs_check(salmon_gpu_set_projectors(gctx, NI, Nps, Mps, Jxyz, Jxx, Jyy, Jzz, Nlma, a_tbl, uV, iuV), &
         'set_projectors')
    allocate(k0(NK_s:NK_e, 3), oc(NB, NK_s:NK_e))
    k0 = kAc0(NK_s:NK_e, :)
    oc = occ(1:NB, NK_s:NK_e)
    call arted_gpu_gs_check(salmon_gpu_set_kpoints(gctx, NK_e - NK_s + 1, NB, k0, oc), 'set_kpoints')
    gready = .true.
  end subroutine
end module arted_gpu_gs_state

subroutine CG_omp(iter_cg_max)
  use Global_Variables, only: NB, NK, NK_s, NK_e, kAc, zu_GS, esp_var, Vloc
  use salmon_parallel, only: nproc_group_tdks
  use salmon_communication, only: comm_summation
  use arted_gpu_gs_state
  implicit none
  integer :: iter_cg_max
  real(8), allocatable :: k(:,:), var(:,:), esp_var_l(:,:)
  call arted_gpu_gs_setup()
  ngs_calls = ngs_calls + 1
  allocate(k(NK_s:NK_e, 3), var(NB, NK_s:NK_e), esp_var_l(NB, NK))
  k = kAc(NK_s:NK_e, :)
  call arted_gpu_gs_check(salmon_gpu_set_potential(gctx, Vloc), 'set_potential')
  call arted_gpu_gs_check(salmon_gpu_put_orbitals(gctx, zu_GS), 'put_orbitals')
  call arted_gpu_gs_check(salmon_gpu_cg(gctx, k, iter_cg_max, SALMON_PC_NONE, var, c_null_ptr), 'cg')
  call arted_gpu_gs_check(salmon_gpu_get_orbitals(gctx, zu_GS), 'get_orbitals')
  ! the reference sums the k-local residuals over the processes (CG.f90:131-133)
  esp_var_l = 0.d0
  esp_var_l(:, NK_s:NK_e) = var
  call comm_summation(esp_var_l, esp_var, NB * NK, nproc_group_tdks)
end subroutine CG_omp

subroutine Gram_Schmidt
  use Global_Variables, only: zu_GS
  use arted_gpu_gs_state
  implicit none
  call arted_gpu_gs_setup()
  ngs_calls = ngs_calls + 1
  call arted_gpu_gs_check(salmon_gpu_put_orbitals(gctx, zu_GS), 'put_orbitals')
  call arted_gpu_gs_check(salmon_gpu_gram_schmidt(gctx), 'gram_schmidt')
  call arted_gpu_gs_check(salmon_gpu_get_orbitals(gctx, zu_GS), 'get_orbitals')
end subroutine Gram_Schmidt
