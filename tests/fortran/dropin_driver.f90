!-----------------------------------------------------------------------
! dropin_driver.f90 -- test driver for the Fortran drop-ins.
! Same module setup as a host program of the reference solvers, then a
! plain `call pcgsip(fi, ip)` etc.: the linker resolves those names to the
! drop-ins in cap_dropin.f90 (GPU library), not to the reference files.
! usage: dropin_driver <sipsol|pcgsip|iccg|cgstab_sip> <in.bin> <out.bin> <log.txt>
! File formats as oracle/ref_driver.f90.
!-----------------------------------------------------------------------
program dropin_driver
  use types
  use parameters
  use indexes
  use coef
  use coefb
  implicit none
  character(len=256) :: mode, fin, fout, flog
  integer :: ni_, nj_, nk_, nsw_, n
  real(8) :: alfa_, sor_
  real(8), allocatable :: fi(:)

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
  ltest = .false.
  icorr = 1
  ipcorr = 1
  allocate(ae(n),aw(n),an(n),as(n),at(n),ab(n),ap(n),su(n),res(n))
  allocate(bb(n),bs(n),bw(n),bp(n),bn(n),be(n),bt(n))
  allocate(fi(n))
  read(10) fi, ae, aw, an, as, at, ab, ap, su, res
  read(10) bb, bs, bw, bp, bn, be, bt
  close(10)
  open(unit=66, file=trim(flog), status='replace')
  select case (trim(mode))
  case ('sipsol')
    call sipsol(fi, ip)
  case ('pcgsip')
    call pcgsip(fi, ip)
  case ('iccg')
    call iccg(fi, ip)
  case ('cgstab_sip')
    call cgstab_sip(fi, ip)
  case default
    stop 2
  end select
  close(66)
  open(unit=11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) fi, res, bb, bs, bw, bp, bn, be, bt
  write(11) resor(ip)
  close(11)
end program dropin_driver
