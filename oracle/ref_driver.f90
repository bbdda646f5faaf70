!-----------------------------------------------------------------------
! ref_driver.f90 -- test driver that calls the reference's own hot-path
! routines (compiled from the reference sources by oracle/Makefile into
! oracle/_ref/) on inputs read from a binary file, and writes their outputs.
!
! TEST INFRASTRUCTURE ONLY (golden-vector generator for tests/golden/).
!
! It provides no stand-in for any reference routine, header or library: it
! only sets the module state a caller of these routines sets (extents via the
! reference's own setind, allocation of the module arrays, sor/nsw/alfa) and
! calls them.  Routines whose reference files need MPI (pcgsip, iccg call
! global_sum/exchange) are not built.
!
! usage: ref_driver <mode> <in.bin> <out.bin> <log.txt>
!   mode = sipsol | cgstab_sip | grad_gauss | bpres
! in.bin  : int32 ni,nj,nk,nsw ; real64 alfa,sor ; then mode-specific arrays
! out.bin : mode-specific arrays, then real64 resor(ip) for the solvers
!-----------------------------------------------------------------------
program ref_driver
  use types
  use parameters
  use indexes
  use coef
  use coefb
  use geometry
  implicit none
  character(len=256) :: mode, fin, fout, flog
  integer :: ni_, nj_, nk_, nsw_, n
  real(8) :: alfa_, sor_
  real(8), allocatable :: fi(:), d1(:), d2(:), d3(:)

  call get_command_argument(1, mode)
  call get_command_argument(2, fin)
  call get_command_argument(3, fout)
  call get_command_argument(4, flog)

  open(unit=10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) ni_, nj_, nk_, nsw_
  read(10) alfa_, sor_

  ni = ni_; nj = nj_; nk = nk_
  n = ni*nj*nk
  nxyza = n
  nxyz = n
  allocate(li(ni), lk(nk))
  call setind

  alfa = alfa_
  sor(ip) = sor_
  nsw(ip) = nsw_
  resor(ip) = 0.0d0
  ltest = .false.
  icorr = 1
  ipcorr = 1

  allocate(ae(n),aw(n),an(n),as(n),at(n),ab(n),ap(n),su(n),res(n))
  allocate(bb(n),bs(n),bw(n),bp(n),bn(n),be(n),bt(n))
  allocate(fx(n),fy(n),fz(n),vol(n))
  allocate(ar1x(n),ar1y(n),ar1z(n),ar2x(n),ar2y(n),ar2z(n),ar3x(n),ar3y(n),ar3z(n))
  allocate(fi(n), d1(n), d2(n), d3(n))

  open(unit=66, file=trim(flog), status='replace')
  open(unit=11, file=trim(fout), access='stream', form='unformatted', status='replace')

  select case (trim(mode))
  case ('sipsol', 'cgstab_sip')
    read(10) fi, ae, aw, an, as, at, ab, ap, su, res
    read(10) bb, bs, bw, bp, bn, be, bt
    if (trim(mode) == 'sipsol') then
      call sipsol(fi, ip)
    else
      call cgstab_sip(fi, ip)
    end if
    write(11) fi, res, bb, bs, bw, bp, bn, be, bt
    write(11) resor(ip)
  case ('grad_gauss')
    read(10) fi, fx, fy, fz, ar1x, ar1y, ar1z, ar2x, ar2y, ar2z, ar3x, ar3y, ar3z, vol
    call grad_gauss(fi, d1, d2, d3)
    write(11) d1, d2, d3
  case ('bpres')
    read(10) fi, fx, fy, fz
    call bpres(fi)
    write(11) fi
  case default
    print *, 'unknown mode ', trim(mode)
    stop 2
  end select
  close(10)
  close(11)
  close(66)
end program ref_driver
