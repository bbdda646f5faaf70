!-----------------------------------------------------------------------
! cap_dropin.f90 -- drop-in replacements for the reference's linear
! solvers with their exact interfaces (implicit-interface external
! subroutines taking (fi, ifi)):
!   sipsol     sipsol.f90:3       pcgsip      pcg-sip.f90:3
!   iccg       iccg.f90:3         cgstab_sip  cgstab_sip.f90:1
! Build: compile this file together with the host program's own module
! file (the one defining modules types, parameters, indexes, coef, coefb,
! title_mod) in place of the four solver files, and link libcapamd.so.
! Like the originals they read ae..ap, su, res, bb..bt, alfa, sor(ifi) and
! nsw(ifi) from the modules; they write fi, res, bb..bt and resor(ifi) and
! print the same one-line solver report on unit 66.
!-----------------------------------------------------------------------
subroutine cap_check(rc, what)
  use iso_c_binding
  use cap_hip
  implicit none
  integer(c_int), intent(in) :: rc
  character(len=*), intent(in) :: what
  if (rc < 0) then
    write(*, '(3a,i0)') 'cap_hip: ', what, ' failed with status ', rc
    stop 3
  end if
end subroutine cap_check

subroutine sipsol(fi, ifi)
  use types; use parameters; use indexes; use coef; use coefb; use title_mod
  use iso_c_binding; use cap_hip
  implicit none
  integer, intent(in) :: ifi
  real(prec), dimension(nxyza) :: fi
  type(cap_coef) :: c
  type(cap_report) :: rep
  integer(c_int) :: rc
  call cap_pack(c, ae, aw, an, as, at, ab, ap, su, res, bb, bs, bw, bp, bn, be, bt)
  rc = cap_sipsol(cap_context(ni, nj, nk), cap_addr(fi), c, alfa, sor(ifi), &
                  int(nsw(ifi), c_int), rep)
  call cap_check(rc, 'sipsol')
  resor(ifi) = rep%resor
  write(66,'(3a,1PE10.3,a,1PE10.3,a,I0)') &
  'SIP:  Solving for ',trim(chvarSolver(IFI)), &
  ', Initial residual = ',RESOR(IFI),', Final residual = ',rep%resl,', No Iterations ',rep%iters
end subroutine sipsol

subroutine pcgsip(fi, ifi)
  use types; use parameters; use indexes; use coef; use coefb; use title_mod
  use iso_c_binding; use cap_hip
  implicit none
  integer, intent(in) :: ifi
  real(prec), dimension(nxyza) :: fi
  type(cap_coef) :: c
  type(cap_report) :: rep
  integer(c_int) :: rc
  call cap_pack(c, ae, aw, an, as, at, ab, ap, su, res, bb, bs, bw, bp, bn, be, bt)
  rc = cap_pcgsip(cap_context(ni, nj, nk), cap_addr(fi), c, alfa, sor(ifi), &
                  int(nsw(ifi), c_int), rep)
  call cap_check(rc, 'pcgsip')
  resor(ifi) = rep%resor
  write(66,'(3a,1pe10.3,a,1pe10.3,a,i0)') &
  'PCG(SIP):  Solving for ',trim(chvarsolver(ifi)), &
  ', Initial residual = ',rep%resor,', fFinal residual = ',rep%resl,', No Iterations ',rep%iters
end subroutine pcgsip

subroutine iccg(fi, ifi)
  use types; use parameters; use indexes; use coef; use coefb; use title_mod
  use iso_c_binding; use cap_hip
  implicit none
  integer, intent(in) :: ifi
  real(prec), dimension(nxyza) :: fi
  type(cap_coef) :: c
  type(cap_report) :: rep
  integer(c_int) :: rc
  ! iccg keeps no SIP factors: only the coef arrays and res are handed over
  c%ae = cap_addr(ae); c%aw = cap_addr(aw); c%an = cap_addr(an); c%as = cap_addr(as)
  c%at = cap_addr(at); c%ab = cap_addr(ab); c%ap = cap_addr(ap); c%su = cap_addr(su)
  c%res = cap_addr(res)
  rc = cap_iccg(cap_context(ni, nj, nk), cap_addr(fi), c, sor(ifi), int(nsw(ifi), c_int), rep)
  call cap_check(rc, 'iccg')
  resor(ifi) = rep%resor
  write(66,'(3a,1PE10.3,a,1PE10.3,a,I0)') &
  'PCG(IC0):  Solving for ',trim(chvarSolver(IFI)), &
  ', Initial residual = ',rep%resor,', Final residual = ',rep%resl,', No Iterations ',rep%iters
end subroutine iccg

subroutine cgstab_sip(fi, ifi)
  use types; use parameters; use indexes; use coef; use coefb; use title_mod
  use iso_c_binding; use cap_hip
  implicit none
  integer, intent(in) :: ifi
  real(prec), dimension(nxyza) :: fi
  type(cap_coef) :: c
  type(cap_report) :: rep
  integer(c_int) :: rc
  call cap_pack(c, ae, aw, an, as, at, ab, ap, su, res, bb, bs, bw, bp, bn, be, bt)
  rc = cap_cgstab_sip(cap_context(ni, nj, nk), cap_addr(fi), c, alfa, sor(ifi), &
                      int(nsw(ifi), c_int), rep)
  call cap_check(rc, 'cgstab_sip')
  resor(ifi) = rep%resor
  write(66,'(3a,i1,a,i1,2a,1PE10.3,a,1PE10.3,a,I0)') &
  'BiCGStab(SIP):  Solving for ',trim(chvarSolver(IFI)),'(icorr=',icorr,',npcor=',ipcorr,')', &
  ', Initial residual = ',rep%resor,', Final residual = ',rep%resl,', No Iterations ',rep%iters
end subroutine cgstab_sip
