! Writes the gs_wfn_k files and rt.data-style rows with the Fortran I/O
! statements SALMON uses (src/ARTED/GS/io_gs_wfn_k.f90:49-68: unformatted
! sequential records of occ(NB,NK) and zu(NL,NB) per k; control_sc.f90:537:
! the row format), from values given in in.bin (stream: nl, nb, nk, nv as
! int32, then zu(nl,nb,nk), occ(nb,nk), v(nv) as real(8)/complex(8)).  Test
! infrastructure for the readers and writers of salmon_amd/formats.py.
program formats_probe
  implicit none
  integer :: nl, nb, nk, nv, ik, i, u
  complex(8), allocatable :: zu(:,:,:)
  real(8), allocatable :: occ(:,:), v(:)
  character(256) :: fname

  open(newunit=u, file='in.bin', access='stream', form='unformatted', status='old')
  read(u) nl, nb, nk, nv
  allocate(zu(nl,nb,nk), occ(nb,nk), v(nv))
  read(u) zu, occ, v
  close(u)

  open(newunit=u, file='gs_wfn_k/occupation', form='unformatted')
  write(u) occ
  close(u)
  do ik = 1, nk
    write(fname, '(A,I7.7,A)') 'gs_wfn_k/wfn_gs_k', ik, '.wfn'
    open(newunit=u, file=trim(fname), form='unformatted')
    write(u) zu(:,:,ik)
    close(u)
  end do

  open(newunit=u, file='rows.txt')
  do i = 1, nv, 5
    write(u, "(F16.8,99(1X,E23.15E3))", advance='no') v(i:min(i+4, nv))
    write(u, *)
  end do
  close(u)
end program formats_probe
