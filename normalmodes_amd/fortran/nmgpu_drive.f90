!-----------------------------------------------------------------------
! A Fortran host driving the device path in the reference's call order:
! setupmatvec (src/mod_matvec.f90:24-95) then pnm_apply_pevsl
! (src/mod_pevsl.f90:54-162).  Reads the unscaled CSR matrices A and B
! (0-based int32 offsets and columns, float64 values) from a stream file:
!   n, nnzA, nnzB (int32); iaA(n+1), jaA(nnzA), aA(nnzA); iaB(n+1), jaB(nnzB), aB(nnzB)
! and the interval [a, b] and B-solve degree from the command line:
!   nmgpu_drive <file> <a> <b> <bdeg>
! Prints the eigenvalues, the reference's residual and the 2-norm residual.
!-----------------------------------------------------------------------
program nmgpu_drive
  use iso_c_binding
  use nmgpu_f90
  implicit none
  character(len=1024) :: fname, arg
  integer(c_int) :: n, nnza, nnzb, ierr, nev, bdeg, rs(2)
  integer(c_int), allocatable :: iaA(:), jaA(:), iaB(:), jaB(:)
  real(c_double), allocatable :: aA(:), aB(:), d(:), eigval(:), eigvec(:), r1(:), r2(:)
  integer(c_int64_t) :: matA, matB, hB, h, chebB, pol
  real(c_double) :: lmin, lmax, blmin, blmax, xintv(4), a, b
  integer :: u, j, mlan, maxit

  call get_command_argument(1, fname)
  call get_command_argument(2, arg); read (arg, *) a
  call get_command_argument(3, arg); read (arg, *) b
  call get_command_argument(4, arg); read (arg, *) bdeg

  open (newunit=u, file=trim(fname), access='stream', form='unformatted', status='old')
  read (u) n, nnza, nnzb
  allocate (iaA(n + 1), jaA(nnza), aA(nnza), iaB(n + 1), jaB(nnzb), aB(nnzb))
  read (u) iaA, jaA, aA, iaB, jaB, aB
  close (u)
  rs = [0, n]

  ! ---- setupmatvec: matrices to HBM, B scaled there, bounds and chebiter of B
  call nmg_check(nmg_parcsr_create(n, n, rs, rs, iaB, jaB, aB, 0, matB), 'parcsr B')
  call nmg_check(nmg_parcsr_create(n, n, rs, rs, iaA, jaA, aA, 0, matA), 'parcsr A')
  allocate (d(n))
  call nmg_check(nmg_csr_diagscale(matB, 1.0d0, d), 'Bdiagscaling')
  call nmg_check(nmg_start(0, hB), 'start B')
  call nmg_check(nmg_setprobsizes(hB, int(n, c_int64_t), int(n, c_int64_t), -1_c_int64_t), 'sizes B')
  call nmg_check(nmg_set_amv(hB, matB), 'amv B')
  call nmg_check(nmg_lanbounds(hB, 1000, 2000, 1.0d-12, blmin, blmax), 'lanbounds B')
  call nmg_check(nmg_finish(hB), 'finish B')
  call nmg_check(nmg_setup_chebiter(blmin, blmax, bdeg, matB, chebB), 'chebiter setup')

  ! ---- pnm_apply_pevsl
  call nmg_check(nmg_start(0, h), 'start')
  call nmg_check(nmg_setprobsizes(h, int(n, c_int64_t), int(n, c_int64_t), -1_c_int64_t), 'sizes')
  call nmg_check(nmg_set_bmv(h, matB), 'bmv')
  call nmg_check(nmg_set_bsol_chebiter(h, 2, chebB), 'bsol')
  call nmg_check(nmg_set_amv_dad(h, matA, d), 'amv')
  call nmg_check(nmg_set_geneig(h), 'geneig')
  call nmg_check(nmg_lanbounds(h, 3000, 5000, 1.0d-5, lmin, lmax), 'lanbounds')
  xintv = [a, b, lmin, lmax]
  if (xintv(1) < 1.0d-10) xintv(1) = lmin
  call nmg_check(nmg_findpol(xintv, 0.8d0, 0.7d0, pol), 'findpol')
  mlan = max(4*802, 2000)
  maxit = 3*mlan
  call nmg_check(nmg_cheblannr(h, xintv, maxit, 1.0d-5, pol), 'cheblannr')
  call nmg_check(nmg_get_nev(h, nev), 'get_nev')
  allocate (eigval(nev), eigvec(int(nev, 8)*n), r1(nev), r2(nev))
  call nmg_check(nmg_copy_result(h, eigval, eigvec, int(n, c_int64_t)), 'copy_result')
  call nmg_check(nmg_residuals(h, r1, r2), 'residuals')
  write (*, '(a,i0)') 'nev ', nev
  do j = 1, nev
    write (*, '(a,i4,3es24.15)') 'eig', j, eigval(j), r1(j), r2(j)
  end do
  call nmg_check(nmg_freepol(pol), 'freepol')
  call nmg_check(nmg_finish(h), 'finish')
  call nmg_check(nmg_chebiter_free(chebB), 'chebiter free')
  call nmg_check(nmg_csr_free(matA), 'free A')
  call nmg_check(nmg_csr_free(matB), 'free B')
end program nmgpu_drive
