!-----------------------------------------------------------------------
! write_grid.f90 -- test writer of the grid file that read_grid.f90:126-161
! reads: Fortran unformatted sequential records in the order of its read
! statements, written by this compiler's own record format, so that the
! C reader (cap_read_grid) is checked against real Fortran records.
! usage: write_grid <in.bin> <grid file>
! in.bin (stream): int32 ni, nj, nk; int32 lbw, lbe (nj*nk), lbs, lbn
! (ni*nk), lbb, lbt (ni*nj); int32 flags(9); real64 x, y, z, xc, yc, zc, vol,
! fx, fy, fz, ar1x .. ar3z (ni*nj*nk each).
!-----------------------------------------------------------------------
program write_grid
  implicit none
  character(len=256) :: fin, fout
  integer :: ni, nj, nk, nij, nik, njk, nijk, q, ijk
  integer, allocatable :: lbw(:), lbe(:), lbs(:), lbn(:), lbb(:), lbt(:)
  integer :: fl(9)
  logical :: low, loe, los, lon, lob, lot, lout, lthermbc, lconbc
  real(8), allocatable :: a(:, :)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(unit=10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) ni, nj, nk
  nij = ni*nj; nik = ni*nk; njk = nj*nk; nijk = nij*nk
  allocate(lbw(njk), lbe(njk), lbs(nik), lbn(nik), lbb(nij), lbt(nij), a(nijk, 19))
  read(10) lbw, lbe, lbs, lbn, lbb, lbt, fl
  read(10) a
  close(10)
  low = fl(1) /= 0; loe = fl(2) /= 0; los = fl(3) /= 0; lon = fl(4) /= 0; lob = fl(5) /= 0
  lot = fl(6) /= 0; lout = fl(7) /= 0; lthermbc = fl(8) /= 0; lconbc = fl(9) /= 0
  open(unit=11, file=trim(fout), form='unformatted', status='replace')
  write(11) ni, nj, nk
  write(11) nij, nik, njk, nijk, 0, 0, 0, 0, &
            (lbw(q), q=1,njk), (lbe(q), q=1,njk), (lbs(q), q=1,nik), (lbn(q), q=1,nik), &
            (lbb(q), q=1,nij), (lbt(q), q=1,nij), &
            low, loe, los, lon, lob, lot, lout, lthermbc, lconbc
  do q = 1, 19
    write(11) (a(ijk, q), ijk=1,nijk)
  end do
  close(11)
end program write_grid
