! Drop-in check: calls the reference's entry points with their Fortran argument
! lists on a 1 x 1 grid, as Test/psmma_testtoeplitz.f90 does on P x Q.
! Toeplitz generator of the reference driver: D(I) = N-I (I /= N), D(N) = N*N,
! i.e. H = N^2 I + S with S(i,j) = i-j, so A*H = N^2 A + (A r) 1' - (A 1) r'.
! Prints the relative error of the whole product and of a Hankel product.
program psmma_fortran_driver
  implicit none
  integer, parameter :: n = 300
  integer :: i, j, desca(9), ldta
  double precision, allocatable :: a(:,:), q(:,:), want(:,:), d(:), work(:), ar(:), a1(:), h(:)
  double precision :: err, nrm, r(n)
  external psmma_toeplitz, psmma_hankel

  allocate(a(n,n), q(n,n), want(n,n), d(2*n-1), work(1), ar(n), a1(n), h(2*n-1))
  ldta = 2*n - 1
  do i = 1, n - 1
    d(i) = dble(n - i)
  end do
  d(n) = dble(n) * dble(n)
  do i = n + 1, ldta
    d(i) = dble(n - i)
  end do
  do j = 1, n
    do i = 1, n
      a(i, j) = sin(dble(i) * 0.37d0 + dble(j) * 1.13d0)
    end do
    r(j) = dble(j)
  end do
  desca = (/ 1, 0, n, n, 64, 64, 0, 0, n /)
  q = 0.0d0
  call psmma_toeplitz(n, n, 1.0d0, a, 1, 1, desca, d, ldta, 0.0d0, q, 1, 1, desca, work, .false.)
  ar = matmul(a, r)
  a1 = sum(a, dim=2)
  do j = 1, n
    do i = 1, n
      want(i, j) = dble(n) * dble(n) * a(i, j) + ar(i) - a1(i) * r(j)
    end do
  end do
  err = sqrt(sum((q - want)**2)) / sqrt(sum(want**2))
  write(*, '(A,ES12.4)') 'toeplitz_rel_err ', err

  ! Hankel H(i,j) = h(i+j-1) against an explicit product
  do i = 1, 2*n - 1
    h(i) = cos(dble(i) * 0.71d0)
  end do
  do j = 1, n
    do i = 1, n
      want(i, j) = 0.0d0
    end do
  end do
  do j = 1, n
    do i = 1, n
      want(:, j) = want(:, j) + a(:, i) * h(i + j - 1)
    end do
  end do
  q = 0.0d0
  call psmma_hankel(n, n, 1.0d0, a, 1, 1, desca, h, ldta, 0.0d0, q, 1, 1, desca, work, .true.)
  nrm = sqrt(sum(want**2))
  err = sqrt(sum((q - want)**2)) / nrm
  write(*, '(A,ES12.4)') 'hankel_rel_err ', err
end program psmma_fortran_driver
