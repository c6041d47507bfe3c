!-----------------------------------------------------------------------
! nmgpu_f90: ISO_C_BINDING interfaces to the nmgpu C-ABI (include/nmgpu.h)
! for a Fortran host.  Each interface replaces one pEVSL F90 call that
! NormalModes makes (src/mod_pevsl.f90, src/mod_matvec.f90); scalars are
! passed by value, arrays by reference, handles are integer(c_int64_t)
! (the reference keeps pEVSL handles in integer*8,
! src/mod_cg_datatype.f90:92-96).  Every function returns 0 or a negative
! status; nmg_errmsg() returns the message of the last failure.
!-----------------------------------------------------------------------
module nmgpu_f90
  use iso_c_binding
  implicit none
  private :: c_strlen

  integer(c_int), parameter :: NMG_FP64 = 0, NMG_FP32_FILTER = 1

  interface
    function c_strlen(s) bind(C, name="strlen") result(n)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: n
    end function

    function nmg_last_error() bind(C, name="nmg_last_error") result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function

    ! pEVSL_Start_F90 (src/mod_pevsl.f90:54)
    integer(c_int) function nmg_start(comm, h) bind(C, name="nmg_start")
      import :: c_int, c_int64_t
      integer(c_int), value :: comm
      integer(c_int64_t), intent(out) :: h
    end function
    ! PEVSL_FINISH_F90 (src/mod_pevsl.f90:220)
    integer(c_int) function nmg_finish(h) bind(C, name="nmg_finish")
      import :: c_int, c_int64_t
      integer(c_int64_t), value :: h
    end function
    ! PEVSL_SETPROBSIZES_F90 (src/mod_pevsl.f90:57)
    integer(c_int) function nmg_setprobsizes(h, n_glob, n_loc, nfirst) bind(C, name="nmg_setprobsizes")
      import :: c_int, c_int64_t
      integer(c_int64_t), value :: h, n_glob, n_loc, nfirst
    end function
    ! PEVSL_PARCSRCREATE_F90 (src/mod_matvec.f90:69-71)
    integer(c_int) function nmg_parcsr_create(nrow_g, ncol_g, row_starts, col_starts, ia, ja, a, comm, mat) &
        bind(C, name="nmg_parcsr_create")
      import :: c_int, c_int64_t, c_double
      integer(c_int), value :: nrow_g, ncol_g, comm
      integer(c_int), intent(in) :: row_starts(*), col_starts(*), ia(*), ja(*)
      real(c_double), intent(in) :: a(*)
      integer(c_int64_t), intent(out) :: mat
    end function
    integer(c_int) function nmg_csr_free(mat) bind(C, name="nmg_csr_free")
      import :: c_int, c_int64_t
      integer(c_int64_t), value :: mat
    end function
    ! pevsl_parcsrmatvec_f90 (src/mod_matvec.f90:453)
    integer(c_int) function nmg_parcsr_matvec(x, y, mat) bind(C, name="nmg_parcsr_matvec")
      import :: c_int, c_int64_t, c_double
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(out) :: y(*)
      integer(c_int64_t), value :: mat
    end function
    ! Bdiagscaling (src/mod_matvec.f90:252-342); sign -1 for Apdiagscaling
    integer(c_int) function nmg_csr_diagscale(mat, sign, d) bind(C, name="nmg_csr_diagscale")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: mat
      real(c_double), value :: sign
      real(c_double), intent(out) :: d(*)
    end function
    ! PEVSL_SETAMV_F90(h, sparseAV, mymatvec) (src/mod_pevsl.f90:77)
    integer(c_int) function nmg_set_amv_dad(h, matA, d) bind(C, name="nmg_set_amv_dad")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: h, matA
      real(c_double), intent(in) :: d(*)
    end function
    ! PEVSL_SETAMV_F90(h, sparseBV, mymatvec) for the bounds of B (src/mod_matvec.f90:80)
    integer(c_int) function nmg_set_amv(h, mat) bind(C, name="nmg_set_amv")
      import :: c_int, c_int64_t
      integer(c_int64_t), value :: h, mat
    end function
    ! PEVSL_SETBMV_F90 (src/mod_pevsl.f90:69)
    integer(c_int) function nmg_set_bmv(h, matB) bind(C, name="nmg_set_bmv")
      import :: c_int, c_int64_t
      integer(c_int64_t), value :: h, matB
    end function
    ! PEVSL_SETBSOL_CHEBITER_F90 (src/mod_pevsl.f90:72-73)
    integer(c_int) function nmg_set_bsol_chebiter(h, chebtype, cheb) bind(C, name="nmg_set_bsol_chebiter")
      import :: c_int, c_int64_t
      integer(c_int64_t), value :: h, cheb
      integer(c_int), value :: chebtype
    end function
    ! PEVSL_SET_GENEIG_F90 (src/mod_pevsl.f90:82)
    integer(c_int) function nmg_set_geneig(h) bind(C, name="nmg_set_geneig")
      import :: c_int, c_int64_t
      integer(c_int64_t), value :: h
    end function
    integer(c_int) function nmg_set_precision(h, mode) bind(C, name="nmg_set_precision")
      import :: c_int, c_int64_t
      integer(c_int64_t), value :: h
      integer(c_int), value :: mode
    end function
    ! pEVSL_SETUP_CHEBITER_F90 (src/mod_matvec.f90:93)
    integer(c_int) function nmg_setup_chebiter(lmin, lmax, deg, mat, cheb) bind(C, name="nmg_setup_chebiter")
      import :: c_int, c_int64_t, c_double
      real(c_double), value :: lmin, lmax
      integer(c_int), value :: deg
      integer(c_int64_t), value :: mat
      integer(c_int64_t), intent(out) :: cheb
    end function
    ! pevsl_chebiter_f90 (src/mod_matvec.f90:480)
    integer(c_int) function nmg_chebiter(chebtype, b, x, cheb) bind(C, name="nmg_chebiter")
      import :: c_int, c_int64_t, c_double
      integer(c_int), value :: chebtype
      real(c_double), intent(in) :: b(*)
      real(c_double), intent(out) :: x(*)
      integer(c_int64_t), value :: cheb
    end function
    integer(c_int) function nmg_chebiter_free(cheb) bind(C, name="nmg_chebiter_free")
      import :: c_int, c_int64_t
      integer(c_int64_t), value :: cheb
    end function
    ! pEVSL_LANBOUNDS_F90 (src/mod_pevsl.f90:84)
    integer(c_int) function nmg_lanbounds(h, mlan, maxit, tol, lmin, lmax) bind(C, name="nmg_lanbounds")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: h
      integer(c_int), value :: mlan, maxit
      real(c_double), value :: tol
      real(c_double), intent(out) :: lmin, lmax
    end function
    ! pEVSL_FINDPOL_F90 (src/mod_pevsl.f90:115)
    integer(c_int) function nmg_findpol(xintv, thr_int, thr_ext, pol) bind(C, name="nmg_findpol")
      import :: c_int, c_int64_t, c_double
      real(c_double), intent(in) :: xintv(4)
      real(c_double), value :: thr_int, thr_ext
      integer(c_int64_t), intent(out) :: pol
    end function
    ! pEVSL_FREEPOL_F90 (src/mod_pevsl.f90:215)
    integer(c_int) function nmg_freepol(pol) bind(C, name="nmg_freepol")
      import :: c_int, c_int64_t
      integer(c_int64_t), value :: pol
    end function
    ! pEVSL_CHEBLANNR_F90 (src/mod_pevsl.f90:122)
    integer(c_int) function nmg_cheblannr(h, xintv, maxit, tol, pol) bind(C, name="nmg_cheblannr")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: h, pol
      real(c_double), intent(in) :: xintv(4)
      integer(c_int), value :: maxit
      real(c_double), value :: tol
    end function
    ! pEVSL_GET_NEV_F90 (src/mod_pevsl.f90:124)
    integer(c_int) function nmg_get_nev(h, nev) bind(C, name="nmg_get_nev")
      import :: c_int, c_int64_t
      integer(c_int64_t), value :: h
      integer(c_int), intent(out) :: nev
    end function
    ! pEVSL_COPY_RESULT_F90 (src/mod_pevsl.f90:130)
    integer(c_int) function nmg_copy_result(h, val, vec, n_loc) bind(C, name="nmg_copy_result")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: h, n_loc
      real(c_double), intent(out) :: val(*), vec(*)
    end function
    ! residual check of src/mod_pevsl.f90:137-162, on the device
    integer(c_int) function nmg_residuals(h, res_ref, res2) bind(C, name="nmg_residuals")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: h
      real(c_double), intent(out) :: res_ref(*), res2(*)
    end function
    ! sparsefsAV (src/mod_matvec.f90:498-520): D (Ad - E Ap^-1 ET) D
    integer(c_int) function nmg_set_amv_fs(h, matAd, matE, matET, chebAp, d, dp) bind(C, name="nmg_set_amv_fs")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: h, matAd, matE, matET, chebAp
      real(c_double), intent(in) :: d(*), dp(*)
    end function
    ! kernel polynomial density of states (pEVSL kpmdos); intv = (a, b, lmin, lmax)
    integer(c_int) function nmg_kpmdos(h, mdeg, nvec, intv, mu, ecnt) bind(C, name="nmg_kpmdos")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: h
      integer(c_int), value :: mdeg, nvec
      real(c_double), intent(in) :: intv(4)
      real(c_double), intent(out) :: mu(*), ecnt
    end function
    ! equal-count slice boundaries (pEVSL spslicer)
    integer(c_int) function nmg_spslicer(sli, mu, mdeg, intv, n_slices) bind(C, name="nmg_spslicer")
      import :: c_int, c_double
      real(c_double), intent(out) :: sli(*)
      real(c_double), intent(in) :: mu(*), intv(4)
      integer(c_int), value :: mdeg, n_slices
    end function
    ! spectrum slices in lockstep (NSLICES = 1 in src/mod_pevsl.f90:47)
    integer(c_int) function nmg_solve_slices(h, n_slices, sli, lmin, lmax, thr_int, thr_ext, maxit, tol) &
        bind(C, name="nmg_solve_slices")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: h
      integer(c_int), value :: n_slices, maxit
      real(c_double), intent(in) :: sli(*)
      real(c_double), value :: lmin, lmax, thr_int, thr_ext, tol
    end function
    integer(c_int) function nmg_slice_counts(h, counts) bind(C, name="nmg_slice_counts")
      import :: c_int, c_int64_t
      integer(c_int64_t), value :: h
      integer(c_int), intent(out) :: counts(*)
    end function
    ! adjoint: per-element d lambda / d (rho, lambda, mu), JOB = 1
    integer(c_int) function nmg_backward(h, n_elem, np, elems, detJ, invJ, refM, refD, job, d_rho, d_lam, d_mu) &
        bind(C, name="nmg_backward")
      import :: c_int, c_int32_t, c_int64_t, c_double
      integer(c_int64_t), value :: h, n_elem
      integer(c_int), value :: np, job
      integer(c_int32_t), intent(in) :: elems(*)
      real(c_double), intent(in) :: detJ(*), invJ(*), refM(*), refD(*)
      real(c_double), intent(out) :: d_rho(*), d_lam(*), d_mu(*)
    end function
    ! the same with reference gravity (JOB = 2): nodal g0 and element gradients
    integer(c_int) function nmg_backward_grav(h, n_elem, np, elems, detJ, invJ, refM, refD, gnode, gsym, &
        d_rho, d_lam, d_mu) bind(C, name="nmg_backward_grav")
      import :: c_int, c_int32_t, c_int64_t, c_double
      integer(c_int64_t), value :: h, n_elem
      integer(c_int), value :: np
      integer(c_int32_t), intent(in) :: elems(*)
      real(c_double), intent(in) :: detJ(*), invJ(*), refM(*), refD(*), gnode(*), gsym(*)
      real(c_double), intent(out) :: d_rho(*), d_lam(*), d_mu(*)
    end function
    ! device producer of a fluid-solid body (CGFSE3D_ISO,
    ! src/mod_cg_create_matrix.f90:103-977): Ad, B, E, ET, Ap on the layout of
    ! matrixstruct_general (:1458-2033), 0-based ids and row offsets; pass
    ! c_null_ptr for g without gravity and for stats when not wanted
    integer(c_int) function nmg_assemble_fs(n_nodes, n_elem, np, nfp, nodes, elems, vslot, lam, mu, rho, g, &
        efluid, neigh, purefluid, refM, refD, massF, fmask, adj_ptr, adj, vstat, vstt, pstt, &
        ad_ptr, ad_col, ad_val, b_ptr, b_col, b_val, e_ptr, e_col, e_val, et_ptr, et_col, et_val, &
        ap_ptr, ap_col, ap_val, stats) bind(C, name="nmg_assemble_fs")
      import :: c_int, c_int8_t, c_int32_t, c_int64_t, c_double, c_ptr
      integer(c_int64_t), value :: n_nodes, n_elem
      integer(c_int), value :: np, nfp, purefluid
      real(c_double), intent(in) :: nodes(*), lam(*), mu(*), rho(*), refM(*), refD(*), massF(*)
      type(c_ptr), value :: g, stats
      integer(c_int32_t), intent(in) :: elems(*), vslot(*), neigh(*), fmask(*), adj(*)
      integer(c_int8_t), intent(in) :: efluid(*), vstat(*)
      integer(c_int64_t), intent(in) :: adj_ptr(*), vstt(*), pstt(*), ad_ptr(*), b_ptr(*), e_ptr(*), et_ptr(*), &
          ap_ptr(*)
      integer(c_int32_t), intent(out) :: ad_col(*), b_col(*), e_col(*), et_col(*), ap_col(*)
      real(c_double), intent(out) :: ad_val(*), b_val(*), e_val(*), et_val(*), ap_val(*)
    end function
    ! adjoint of a fluid-solid body for k pairs: eig(k), U(n_dof, k) physical
    ! and B-normalised, P(n_p, k) = -Ap^-1 ET U; outputs (n_elem, k)
    integer(c_int) function nmg_backward_fs(n_nodes, n_elem, np, nfp, nodes, elems, vslot, lam, rho, g, &
        efluid, neigh, purefluid, refM, refD, massF, fmask, vstat, vstt, pstt, k, eig, U, P, &
        d_rho, d_lam, d_mu) bind(C, name="nmg_backward_fs")
      import :: c_int, c_int8_t, c_int32_t, c_int64_t, c_double, c_ptr
      integer(c_int64_t), value :: n_nodes, n_elem
      integer(c_int), value :: np, nfp, purefluid, k
      real(c_double), intent(in) :: nodes(*), lam(*), rho(*), refM(*), refD(*), massF(*), eig(*), U(*), P(*)
      type(c_ptr), value :: g
      integer(c_int32_t), intent(in) :: elems(*), vslot(*), neigh(*), fmask(*)
      integer(c_int8_t), intent(in) :: efluid(*), vstat(*)
      integer(c_int64_t), intent(in) :: vstt(*), pstt(*)
      real(c_double), intent(out) :: d_rho(*), d_lam(*), d_mu(*)
    end function
  end interface

contains

  ! message of the last failed call
  function nmg_errmsg() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: i, n
    p = nmg_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, s, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = s(i)
    end do
  end function

  ! stop with the library's message on a failed call (the reference stops
  ! on fatal conditions, e.g. src/mod_utility.f90:1011-1013)
  subroutine nmg_check(ierr, what)
    integer(c_int), intent(in) :: ierr
    character(len=*), intent(in) :: what
    if (ierr /= 0) then
      write (*, '(a,a,i0,a,a)') trim(what), ': status ', ierr, ': ', nmg_errmsg()
      stop 1
    end if
  end subroutine

end module nmgpu_f90
