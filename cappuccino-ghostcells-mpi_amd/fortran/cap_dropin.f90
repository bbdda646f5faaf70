!-----------------------------------------------------------------------
! cap_dropin.f90 -- drop-in replacements for the reference's linear
! solvers with their exact interfaces (implicit-interface external
! subroutines taking (fi, ifi)):
!   sipsol     sipsol.f90:3       pcgsip      pcg-sip.f90:3
!   iccg       iccg.f90:3         cgstab_sip  cgstab_sip.f90:1
! and for the p' stencils, with the reference's argument lists:
!   grad_gauss  GRAD_GAUSS.f90:1   bpres  bpres.f90:3   fluxmass  fluxmass2.f90:3
!   assemble_pressure_eq  assemble_pressure_eq.f90:3
!   piso_assemble_pressure_eq  PISO_assemble_pressure_eq.f90:3
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

! module geometry/variables/coef/gradients arrays -> cap_fields
subroutine cap_fields_of_modules(f)
  use types; use geometry; use variables; use coef; use gradients
  use iso_c_binding; use cap_hip
  implicit none
  type(cap_fields), intent(out) :: f
  f%x = cap_addr(x); f%y = cap_addr(y); f%z = cap_addr(z)
  f%xc = cap_addr(xc); f%yc = cap_addr(yc); f%zc = cap_addr(zc); f%vol = cap_addr(vol)
  f%fx = cap_addr(fx); f%fy = cap_addr(fy); f%fz = cap_addr(fz)
  f%ar1x = cap_addr(ar1x); f%ar1y = cap_addr(ar1y); f%ar1z = cap_addr(ar1z)
  f%ar2x = cap_addr(ar2x); f%ar2y = cap_addr(ar2y); f%ar2z = cap_addr(ar2z)
  f%ar3x = cap_addr(ar3x); f%ar3y = cap_addr(ar3y); f%ar3z = cap_addr(ar3z)
  f%den = cap_addr(den); f%apu = cap_addr(apu); f%apv = cap_addr(apv); f%apw = cap_addr(apw)
  f%u = cap_addr(u); f%v = cap_addr(v); f%w = cap_addr(w); f%p = cap_addr(p)
  f%gradu = cap_addr(gradu); f%gradv = cap_addr(gradv); f%gradw = cap_addr(gradw)
  f%gradp = cap_addr(gradp)
end subroutine cap_fields_of_modules

subroutine grad_gauss(u, dudx, dudy, dudz)
  use types; use parameters; use indexes; use geometry
  use iso_c_binding; use cap_hip
  implicit none
  real(prec), dimension(nxyza), intent(in) :: u
  real(prec), dimension(nxyza), intent(inout) :: dudx, dudy, dudz
  type(c_ptr) :: g(13)
  g = [cap_addr(fx), cap_addr(fy), cap_addr(fz), cap_addr(ar1x), cap_addr(ar1y), cap_addr(ar1z), &
       cap_addr(ar2x), cap_addr(ar2y), cap_addr(ar2z), cap_addr(ar3x), cap_addr(ar3y), &
       cap_addr(ar3z), cap_addr(vol)]
  call cap_check(cap_grad_gauss(cap_context(ni, nj, nk), cap_caddr(u), g, cap_addr(dudx), &
                                cap_addr(dudy), cap_addr(dudz)), 'grad_gauss')
end subroutine grad_gauss

subroutine bpres(p)
  use types; use parameters; use indexes; use geometry
  use iso_c_binding; use cap_hip
  implicit none
  real(prec), dimension(nxyza) :: p
  call cap_check(cap_bpres(cap_context(ni, nj, nk), cap_addr(p), cap_addr(fx), cap_addr(fy), &
                           cap_addr(fz)), 'bpres')
end subroutine bpres

subroutine fluxmass(idew, idns, idtb, arvx, arvy, arvz, fif, acfe, acfw, fcf)
  use types; use parameters; use indexes
  use iso_c_binding; use cap_hip
  implicit none
  integer, intent(in) :: idew, idns, idtb
  real(prec), dimension(nxyza) :: arvx, arvy, arvz, fif, acfe, acfw, fcf
  type(cap_fields) :: f
  call cap_fields_of_modules(f)
  call cap_check(cap_fluxmass(cap_context(ni, nj, nk), f, int(idew, c_int), int(idns, c_int), &
                              int(idtb, c_int), cap_addr(arvx), cap_addr(arvy), cap_addr(arvz), &
                              cap_addr(fif), cap_addr(acfe), cap_addr(acfw), cap_addr(fcf)), &
                 'fluxmass')
end subroutine fluxmass

subroutine assemble_pressure_eq(idew, idns, idtb, arvx, arvy, arvz, fif, acfe, acfw, fcf)
  use types; use parameters; use indexes; use coef
  use iso_c_binding; use cap_hip
  implicit none
  integer, intent(in) :: idew, idns, idtb
  real(prec), dimension(nxyza) :: arvx, arvy, arvz, fif, acfe, acfw, fcf
  type(cap_fields) :: f
  call cap_fields_of_modules(f)
  call cap_check(cap_assemble_pressure_eq(cap_context(ni, nj, nk), f, int(idew, c_int), &
                 int(idns, c_int), int(idtb, c_int), cap_addr(arvx), cap_addr(arvy), &
                 cap_addr(arvz), cap_addr(fif), cap_addr(acfe), cap_addr(acfw), cap_addr(fcf), &
                 cap_addr(su), cap_addr(sv), cap_addr(sw)), 'assemble_pressure_eq')
end subroutine assemble_pressure_eq

subroutine piso_assemble_pressure_eq(idew, idns, idtb, arvx, arvy, arvz, fif, acfe, acfw, fcf, &
                                     ipc, bp)
  use types; use parameters; use indexes; use coef
  use iso_c_binding; use cap_hip
  implicit none
  integer, intent(in) :: idew, idns, idtb, ipc
  real(prec), dimension(nxyza) :: arvx, arvy, arvz, fif, acfe, acfw, fcf, bp
  type(cap_fields) :: f
  call cap_fields_of_modules(f)
  call cap_check(cap_piso_assemble_pressure_eq(cap_context(ni, nj, nk), f, int(idew, c_int), &
                 int(idns, c_int), int(idtb, c_int), cap_addr(arvx), cap_addr(arvy), &
                 cap_addr(arvz), cap_addr(fif), cap_addr(acfe), cap_addr(acfw), cap_addr(fcf), &
                 int(ipc, c_int), cap_addr(bp), cap_addr(su)), 'piso_assemble_pressure_eq')
end subroutine piso_assemble_pressure_eq
