! Explicit ISO_C_BINDING interfaces to the device API of libpsmma_hip.so
! (include/psmma.h).  The Fortran drop-ins (PSMMA_TOEPLITZ, ...) need no
! interface: they are called exactly like the reference's externals.
module psmma_hip
  use iso_c_binding
  implicit none
  integer(c_int), parameter :: PSMMA_PREC_FP64 = 0, PSMMA_PREC_SPLIT_I8 = 1, &
                               PSMMA_PREC_SPLIT_BF16 = 2
  interface
    integer(c_int) function psmma_toeplitz_dev(m, n, k, alpha, a, lda, g, lg, korder, r0, c0, &
                                               beta, c, ldc, prec, slices, stream) &
        bind(C, name="psmma_toeplitz_dev")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: m, n, k, lda, lg, korder, r0, c0, ldc, prec, slices
      real(c_double), value :: alpha, beta
      type(c_ptr), value :: a, g, c, stream
    end function
    integer(c_int) function psmma_hankel_dev(m, n, k, alpha, a, lda, g, lg, r0, c0, beta, c, ldc, &
                                             prec, slices, stream) bind(C, name="psmma_hankel_dev")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: m, n, k, lda, lg, r0, c0, ldc, prec, slices
      real(c_double), value :: alpha, beta
      type(c_ptr), value :: a, g, c, stream
    end function
    integer(c_int) function psmma_cauchy_dev(m, n, k, alpha, a, lda, d, f, u, v, r0, c0, beta, c, &
                                             ldc, prec, slices, stream) bind(C, name="psmma_cauchy_dev")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: m, n, k, lda, r0, c0, ldc, prec, slices
      real(c_double), value :: alpha, beta
      type(c_ptr), value :: a, d, f, u, v, c, stream
    end function
    integer(c_int) function psmma_cauchyeig_dev(m, n, k, alpha, a, lda, fpole, u, v, difl, difr, &
                                                nglob, r0, c0, beta, c, ldc, prec, slices, stream) &
        bind(C, name="psmma_cauchyeig_dev")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: m, n, k, lda, nglob, r0, c0, ldc, prec, slices
      real(c_double), value :: alpha, beta
      type(c_ptr), value :: a, fpole, u, v, difl, difr, c, stream
    end function
    integer(c_int) function pzmdft_dev(m, n, k, nd, alpha, b, ldb, r0, c0, beta, c, ldc, prec, &
                                       stream) bind(C, name="pzmdft_dev")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: m, n, k, nd, ldb, r0, c0, ldc, prec
      real(c_double), value :: alpha, beta
      type(c_ptr), value :: b, c, stream
    end function
    integer(c_int) function psmma_set_precision(prec, slices) bind(C, name="psmma_set_precision")
      import :: c_int
      integer(c_int), value :: prec, slices
    end function
    integer(c_int) function psmma_last_status() bind(C, name="psmma_last_status")
      import :: c_int
    end function
  end interface
end module psmma_hip
