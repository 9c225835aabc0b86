! Driver for the reference's generator-only RRLU routines and explicit tile
! generators, compiled from the reference's own module sources (oracle/Makefile,
! target ref) into oracle/_ref.  Test infrastructure only: it runs the
! reference's rrluCauchy (SRC/Cauchylowrank.f90:7-134), rrluCauchyEigrow
! (SRC/CauchyHssEig.f90:7-212), Cauchylike (SRC/Cauchylowrank.f90:509-521) or
! CauchylikeEig (SRC/CauchyHssEig.f90:1693-1774) on one block and writes every
! output, so the oracle's restatements and the GPU kernels' generated tiles can be
! compared with the reference itself.
!
! Usage: ref_rrlu <input> <output>, both unformatted stream files.
!   input:  kind (int32: 0 rrluCauchy, 1 rrluCauchyEigrow, 2 Cauchylike,
!           3 CauchylikeEig), M, N (int32), tol (real64, kinds 0 and 1), then
!           kind 0: D(M), F(N), A(M), V(N)
!           kind 1: nglob (int32), FF, DIFL, DIFR (nglob each), DI(N), A(M), V(N),
!                   PL(M), PU(N) (int32, global 1-based indices)
!           kind 2: D(M), F(N), U(M), V(N)
!           kind 3: nglob (int32), FF, DIFL, DIFR (nglob each), U(M), V(N), PL(M), PU(N)
!   output: kinds 0, 1: Rk (int32), PL(M), PU(N) (int32), then
!           kind 0: D(M), F(N), A(M), V(N), U(M), W(min(M, N)) (real64)
!           kind 1: PQ(M), nswap (int32), F(M), A(M), V(N), U(M), W(min(M, N)) (real64)
!           kinds 2, 3: the M x N block, column-major (real64)
program ref_rrlu
  use cauchylowrank, only: rrluCauchy, Cauchylike
  use cauchyhsseig, only: rrluCauchyEigrow, CauchylikeEig
  implicit none
  integer :: kind, m, n, mn, nglob, rk, nflops, nswap, uin, uout, i
  double precision :: tol
  double precision, allocatable :: d(:), f(:), a(:), v(:), u(:), w(:)
  double precision, allocatable :: ff(:), difl(:), difr(:), di(:), blk(:, :)
  integer, allocatable :: pl(:), pu(:), pq(:)
  character(len=1024) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (newunit=uin, file=trim(fin), access='stream', form='unformatted', status='old')
  read (uin) kind, m, n, tol
  mn = min(m, n)
  if (kind >= 2) then
    allocate (u(m), v(n), blk(m, n), pl(m), pu(n))
    if (kind == 2) then
      allocate (d(m), f(n))
      read (uin) d, f, u, v
      call Cauchylike(blk, d, f, u, v, m, n)
    else
      read (uin) nglob
      allocate (ff(nglob), difl(nglob), difr(nglob))
      read (uin) ff, difl, difr, u, v, pl, pu
      call CauchylikeEig(blk, ff, u, v, difl, difr, pl, pu, m, n, nflops)
    end if
    close (uin)
    open (newunit=uout, file=trim(fout), access='stream', form='unformatted', status='replace')
    write (uout) blk
    close (uout)
    stop
  end if
  allocate (a(m), v(n), u(m), w(max(mn, 1)), pl(m), pu(n))
  u = 0.0d0
  w = 0.0d0
  if (kind == 0) then
    allocate (d(m), f(n))
    read (uin) d, f, a, v
    close (uin)
    call rrluCauchy(d, f, a, v, tol, u, w, pl, pu, m, n, rk)
  else
    read (uin) nglob
    allocate (ff(nglob), difl(nglob), difr(nglob), di(n), f(m), pq(m))
    read (uin) ff, difl, difr, di, a, v, pl, pu
    close (uin)
    ! the block's row poles, as Compress_cauchyEig2 forms them (SRC/CauchyHssEig.f90:2680-2682)
    do i = 1, m
      f(i) = ff(pl(i))
    end do
    call rrluCauchyEigrow(f, di, a, v, difl, difr, tol, u, w, ff, pl, pu, m, n, rk, pq, &
                          nflops, nswap)
  end if
  open (newunit=uout, file=trim(fout), access='stream', form='unformatted', status='replace')
  write (uout) rk, pl, pu
  if (kind == 0) then
    write (uout) d, f, a, v, u, w(1:mn)
  else
    write (uout) pq, nswap, f, a, v, u, w(1:mn)
  end if
  close (uout)
end program ref_rrlu
