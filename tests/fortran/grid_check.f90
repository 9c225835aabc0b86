! A caller on a 2 x 2 BLACS grid.  The drop-ins compute one GPU's share, so they
! must refuse a larger grid before touching the device: PSMMA_TOEPLITZ through
! its descriptor's context, PZMDFTL1 / PZMDFTL through COMMON /CONTEXT/ ICTXT
! (SRC/pzmdftL1.f90:178), the only way the reference passes them a context.
! BLACS_GRIDINFO below stands in for a BLACS library and reports a 2 x 2 grid.
subroutine blacs_gridinfo(ictxt, nprow, npcol, myrow, mycol)
  implicit none
  integer ictxt, nprow, npcol, myrow, mycol
  nprow = 2
  npcol = 2
  myrow = 0
  mycol = 0
  if (ictxt /= 7) nprow = -1  ! the context must be the caller's
end subroutine

program grid_check
  use iso_c_binding, only: c_int
  implicit none
  interface
    integer(c_int) function psmma_last_status() bind(C, name="psmma_last_status")
      import c_int
    end function
  end interface
  integer ictxt
  common /context/ ictxt
  integer, parameter :: n = 8
  double precision a(n, n), q(n, n), g(2*n-1), work(1), rwork(1)
  complex*16 b(n, n), c(n, n)
  integer desca(9)
  external psmma_toeplitz, pzmdftl1, pzmdftl

  ictxt = 7
  desca = (/ 1, 7, n, n, 4, 4, 0, 0, n /)
  a = 1.0d0
  g = 1.0d0
  q = 5.0d0
  b = (1.0d0, 0.0d0)
  c = (3.0d0, 0.0d0)
  call psmma_toeplitz(n, n, 1.0d0, a, 1, 1, desca, g, 2*n-1, 0.0d0, q, 1, 1, desca, work, .false.)
  print '(a,i0)', 'toeplitz_status ', psmma_last_status()
  call pzmdftl1(n, n, 4, 4, 1.0d0, b, n, 0.0d0, c, n, 0, 0, work, rwork)
  print '(a,i0)', 'pzmdftl1_status ', psmma_last_status()
  call pzmdftl(n, n, 4, 4, 1.0d0, b, n, 0.0d0, c, n, 0, 0, work, rwork)
  print '(a,i0)', 'pzmdftl_status ', psmma_last_status()
  if (all(q == 5.0d0) .and. all(c == (3.0d0, 0.0d0))) then
    print '(a)', 'outputs_untouched 1'
  else
    print '(a)', 'outputs_untouched 0'
  end if
end program
