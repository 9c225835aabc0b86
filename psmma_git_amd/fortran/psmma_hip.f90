! Explicit ISO_C_BINDING interfaces to the device API of libpsmma_hip.so
! (include/psmma.h).  The Fortran drop-ins (PSMMA_TOEPLITZ, ...) need no
! interface: they are called exactly like the reference's externals.
module psmma_hip
  use iso_c_binding
  implicit none
  integer(c_int), parameter :: PSMMA_PREC_FP64 = 0, PSMMA_PREC_SPLIT_I8 = 1, &
                               PSMMA_PREC_SPLIT_BF16 = 2, PSMMA_PREC_SPLIT_FP32 = 3
  integer(c_int), parameter :: PSMMA_BLR_RRQR = 0, PSMMA_BLR_RRLU = 1
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
                                       slices, stream) bind(C, name="pzmdft_dev")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: m, n, k, nd, ldb, r0, c0, ldc, prec, slices
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
    ! the merge step's generators: secular roots as DIFL / DIFR, Loewner's z, normalisers
    integer(c_int) function psmma_secular_dev(n, d, z, rho, difl, difr, zhat, v, stream) &
        bind(C, name="psmma_secular_dev")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: n
      real(c_double), value :: rho
      type(c_ptr), value :: d, z, difl, difr, zhat, v, stream
    end function
    ! Compress_RRQR on a batch of explicit blocks
    integer(c_int) function psmma_rrqr_batched_dev(nblk, m, n, a, lda, stride_a, tol, kmax, &
                                                   rank, jpvt, q, ldq, stride_q, rpt, ldr, &
                                                   stride_r, stream) &
        bind(C, name="psmma_rrqr_batched_dev")
      import :: c_int, c_long_long, c_double, c_ptr
      integer(c_int), value :: nblk, m, n, lda, kmax, ldq, ldr
      integer(c_long_long), value :: stride_a, stride_q, stride_r
      real(c_double), value :: tol
      type(c_ptr), value :: a, rank, jpvt, q, rpt, stream
    end function
    ! Compress_cauchyEig2 / compress_cauchy on batches of blocks given by global rows / columns
    integer(c_int) function psmma_rrlu_cauchyeig_batched_dev(nblk, m, n, pl, pu, fpole, u, v, &
                                                             difl, difr, korder, tol, kmax, rank, &
                                                             nswap, q, ldq, stride_q, y, ldy, &
                                                             stride_y, stream) &
        bind(C, name="psmma_rrlu_cauchyeig_batched_dev")
      import :: c_int, c_long_long, c_double, c_ptr
      integer(c_int), value :: nblk, m, n, korder, kmax, ldq, ldy
      integer(c_long_long), value :: stride_q, stride_y
      real(c_double), value :: tol
      type(c_ptr), value :: pl, pu, fpole, u, v, difl, difr, rank, nswap, q, y, stream
    end function
    integer(c_int) function psmma_rrlu_cauchy_batched_dev(nblk, m, n, pl, pu, d, f, u, v, nrow, &
                                                          ncol, tol, kmax, rank, x, ldx, stride_x, &
                                                          y, ldy, stride_y, stream) &
        bind(C, name="psmma_rrlu_cauchy_batched_dev")
      import :: c_int, c_long_long, c_double, c_ptr
      integer(c_int), value :: nblk, m, n, nrow, ncol, kmax, ldx, ldy
      integer(c_long_long), value :: stride_x, stride_y
      real(c_double), value :: tol
      type(c_ptr), value :: pl, pu, d, f, u, v, rank, x, y, stream
    end function
    ! compressed products (stats: host array of 5 integer(c_long_long), may be c_null_ptr)
    integer(c_int) function psmma_cauchyeig_blr_dev(m, n, alpha, a, lda, fpole, u, v, difl, difr, &
                                                    leaf, tol, beta, c, ldc, stats, stream) &
        bind(C, name="psmma_cauchyeig_blr_dev")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: m, n, lda, leaf, ldc
      real(c_double), value :: alpha, tol, beta
      type(c_ptr), value :: a, fpole, u, v, difl, difr, c, stats, stream
    end function
    integer(c_int) function psmma_cauchyeig_blr_ex_dev(m, n, alpha, a, lda, fpole, u, v, difl, &
                                                       difr, leaf, tol, compressor, beta, c, ldc, &
                                                       stats, stream) &
        bind(C, name="psmma_cauchyeig_blr_ex_dev")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: m, n, lda, leaf, compressor, ldc
      real(c_double), value :: alpha, tol, beta
      type(c_ptr), value :: a, fpole, u, v, difl, difr, c, stats, stream
    end function
    integer(c_int) function psmma_cauchy_blr_dev(m, n, alpha, a, lda, d, f, u, v, leaf, tol, &
                                                 compressor, beta, c, ldc, stats, stream) &
        bind(C, name="psmma_cauchy_blr_dev")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: m, n, lda, leaf, compressor, ldc
      real(c_double), value :: alpha, tol, beta
      type(c_ptr), value :: a, d, f, u, v, c, stats, stream
    end function
  end interface
end module psmma_hip
