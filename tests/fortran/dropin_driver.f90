!-----------------------------------------------------------------------
! dropin_driver.f90 -- test driver for the Fortran drop-ins.
! Same module setup as a host program of the reference solvers, then a
! plain `call pcgsip(fi, ip)` etc.: the linker resolves those names to the
! drop-ins in cap_dropin.f90 (GPU library), not to the reference files.
! usage: dropin_driver <mode> <in.bin> <out.bin> <log.txt>
!   mode = sipsol | pcgsip | iccg | cgstab_sip | cgstab | pcgjac   (solvers)
!        | fluxmc | fluxuvw | ghost                             (stencils)
! File formats as oracle/ref_driver.f90's modes of the same names.
!-----------------------------------------------------------------------
program dropin_driver
  use types
  use parameters
  use indexes
  use coef
  use coefb
  use geometry
  use variables
  use gradients
  use bc
  use inlet
  implicit none
  character(len=256) :: mode, fin, fout, flog
  integer :: ni_, nj_, nk_, nsw_, n, i, j, k, inp, idew, idns, idtb, dirn, fl
  real(8) :: alfa_, sor_, fmcor
  real(8), allocatable :: fi(:), d1(:), d2(:), d3(:), bsc(:), lbd(:, :), inl(:)

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
  allocate(fi(n), d1(n), d2(n), d3(n), bsc(n))
  allocate(x(n),y(n),z(n),xc(n),yc(n),zc(n),vol(n),fx(n),fy(n),fz(n))
  allocate(ar1x(n),ar1y(n),ar1z(n),ar2x(n),ar2y(n),ar2z(n),ar3x(n),ar3y(n),ar3z(n))
  allocate(den(n),apu(n),apv(n),apw(n),u(n),v(n),w(n),p(n),sv(n),sw(n))
  allocate(te(n),ed(n),t(n),vis(n))
  allocate(gradu(3,n),gradv(3,n),gradw(3,n),gradp(3,n))
  open(unit=66, file=trim(flog), status='replace')
  open(unit=11, file=trim(fout), access='stream', form='unformatted', status='replace')
  select case (trim(mode))
  case ('sipsol', 'pcgsip', 'iccg', 'cgstab_sip', 'cgstab', 'pcgjac')
    read(10) fi, ae, aw, an, as, at, ab, ap, su, res
    read(10) bb, bs, bw, bp, bn, be, bt
    select case (trim(mode))
    case ('sipsol')
      call sipsol(fi, ip)
    case ('pcgsip')
      call pcgsip(fi, ip)
    case ('iccg')
      call iccg(fi, ip)
    case ('cgstab_sip')
      call cgstab_sip(fi, ip)
    case ('cgstab')
      call cgstab(fi, ip)
    case default
      call Diagonally_Preconditioned_CG(fi, ip)
    end select
    write(11) fi, res, bb, bs, bw, bp, bn, be, bt
    write(11) resor(ip)
  case ('fluxmc', 'fluxuvw')
    read(10) x, y, z, xc, yc, zc, vol, fx, fy, fz, ar1x, ar1y, ar1z, ar2x, ar2y, ar2z, &
             ar3x, ar3y, ar3z, den, apu, apv, apw, u, v, w, p, gradu, gradv, gradw, gradp, &
             su, sv, sw, d1, d2, d3, bsc
    if (trim(mode) == 'fluxmc') then
      dirn = nsw_
    else
      read(10) vis, bt, bb
      bp = bsc
      dirn = mod(nsw_, 8)
      fl = nsw_ / 8
      lcds = iand(fl, 1) /= 0; lcds4 = iand(fl, 2) /= 0; lluds = iand(fl, 4) /= 0
      lquds = iand(fl, 8) /= 0; flux_limiter = iand(fl, 16) /= 0; lsmart = iand(fl, 32) /= 0
      lavl = iand(fl, 64) /= 0; lmuscl = iand(fl, 128) /= 0; lumist = iand(fl, 256) /= 0
      lgamma = iand(fl, 512) /= 0
      gds(iu) = alfa_
    end if
    if (dirn == 0) then
      idew = nj; idns = 1; idtb = nij
    else if (dirn == 1) then
      idew = 1; idns = nij; idtb = nj
    else
      idew = nij; idns = nj; idtb = 1
    end if
    if (trim(mode) == 'fluxmc') then
      ! the cell loop of calcp-multiple_correction_SIMPLE.f90:126-158
      d1 = 0.0d0
      do k = 3, nk-2
      do i = 3, ni-2
      do j = 3, nj-2
        inp = lk(k) + li(i) + j
        fmcor = 0.0d0
        if (dirn == 0) call fluxmc(inp, idew, idns, idtb, fx, fmcor)
        if (dirn == 1) call fluxmc(inp, idew, idns, idtb, fy, fmcor)
        if (dirn == 2) call fluxmc(inp, idew, idns, idtb, fz, fmcor)
        d1(inp) = fmcor
      end do
      end do
      end do
      write(11) d1, d2, d3, bsc, su
    else
      if (dirn == 0) call fluxuvw(idew, idns, idtb, ar1x, ar1y, ar1z, fx, d1, d2, d3)
      if (dirn == 1) call fluxuvw(idew, idns, idtb, ar2x, ar2y, ar2z, fy, d1, d2, d3)
      if (dirn == 2) call fluxuvw(idew, idns, idtb, ar3x, ar3y, ar3z, fz, d1, d2, d3)
      write(11) d1, d2, su, sv, sw, bp, bt, bb
    end if
  case ('ghost')
    allocate(lbd(n, 6), inl(n))
    allocate(lbw(nj*nk), lbe(nj*nk), lbs(ni*nk), lbn(ni*nk), lbb(ni*nj), lbt(ni*nj))
    read(10) u, v, w, te, ed, t, vis, fx, fy, fz, ar1x, ar1y, ar1z, ar2x, ar2y, ar2z, &
             ar3x, ar3y, ar3z
    read(10) lbd, inl
    lbw = int(lbd(1:nj*nk, 1)); lbe = int(lbd(1:nj*nk, 2))
    lbs = int(lbd(1:ni*nk, 3)); lbn = int(lbd(1:ni*nk, 4))
    lbb = int(lbd(1:ni*nj, 5)); lbt = int(lbd(1:ni*nj, 6))
    uin = inl(1); vin = inl(2); win = inl(3); tein = inl(4); edin = inl(5); tin = inl(6)
    call update_values_at_ghostcells
    write(11) u, v, w, te, ed, t, vis
  case default
    stop 2
  end select
  close(10)
  close(11)
  close(66)
end program dropin_driver
