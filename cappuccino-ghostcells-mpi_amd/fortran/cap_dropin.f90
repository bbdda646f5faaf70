!-----------------------------------------------------------------------
! cap_dropin.f90 -- drop-in replacements for the reference's linear
! solvers with their exact interfaces (implicit-interface external
! subroutines taking (fi, ifi)):
!   sipsol     sipsol.f90:3       pcgsip      pcg-sip.f90:3
!   iccg       iccg.f90:3         cgstab_sip  cgstab_sip.f90:1
!   cgstab     cgstab.f90:1       Diagonally_Preconditioned_CG  pcg-jacobi.f90:3
! and for the p' stencils, with the reference's argument lists:
!   grad_gauss  GRAD_GAUSS.f90:1   bpres  bpres.f90:3   fluxmass  fluxmass2.f90:3
!   assemble_pressure_eq  assemble_pressure_eq.f90:3
!   piso_assemble_pressure_eq  PISO_assemble_pressure_eq.f90:3
!   fluxuvw  fluxuvw-overrelaxed-correction.f90:3   fluxmc  fluxmc.f90:3
!   update_values_at_ghostcells  update_values_at_ghostcells_nonperiodic.f90:3
! Build: compile this file together with the host program's own module
! file (the one defining modules types, parameters, indexes, coef, coefb,
! title_mod, geometry, variables, gradients, bc, inlet) in place of the
! replaced files, and link libcapamd.so.
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

subroutine cgstab(fi, ifi)
  use types; use parameters; use indexes; use coef; use title_mod
  use iso_c_binding; use cap_hip
  implicit none
  integer, intent(in) :: ifi
  real(prec), dimension(nxyza) :: fi
  type(cap_coef) :: c
  type(cap_report) :: rep
  integer(c_int) :: rc
  ! no SIP factors: only the coef arrays and res are handed over
  c%ae = cap_addr(ae); c%aw = cap_addr(aw); c%an = cap_addr(an); c%as = cap_addr(as)
  c%at = cap_addr(at); c%ab = cap_addr(ab); c%ap = cap_addr(ap); c%su = cap_addr(su)
  c%res = cap_addr(res)
  rc = cap_cgstab(cap_context(ni, nj, nk), cap_addr(fi), c, sor(ifi), int(nsw(ifi), c_int), rep)
  call cap_check(rc, 'cgstab')
  resor(ifi) = rep%resor
  write(66,'(3a,1PE10.3,a,1PE10.3,a,I0)') &
  'BiCGStab(ILU(0)):  Solving for ',trim(chvarSolver(IFI)), &
  ', Initial residual = ',rep%resor,', Final residual = ',rep%resl,', No Iterations ',rep%iters
end subroutine cgstab

subroutine Diagonally_Preconditioned_CG(fi, ifi)
  use types; use parameters; use indexes; use coef; use title_mod
  use iso_c_binding; use cap_hip
  implicit none
  integer, intent(in) :: ifi
  real(prec), dimension(nxyza) :: fi
  type(cap_coef) :: c
  type(cap_report) :: rep
  integer(c_int) :: rc
  c%ae = cap_addr(ae); c%aw = cap_addr(aw); c%an = cap_addr(an); c%as = cap_addr(as)
  c%at = cap_addr(at); c%ab = cap_addr(ab); c%ap = cap_addr(ap); c%su = cap_addr(su)
  c%res = cap_addr(res)
  rc = cap_pcg_jacobi(cap_context(ni, nj, nk), cap_addr(fi), c, sor(ifi), int(nsw(ifi), c_int), &
                      rep)
  call cap_check(rc, 'Diagonally_Preconditioned_CG')
  resor(ifi) = rep%resor
  write(66,'(3a,1PE10.3,a,1PE10.3,a,I0)') &
  'PCG(Jacobi):  Solving for ',trim(chvarSolver(IFI)), &
  ', Initial residual = ',rep%resor,', Final residual = ',rep%resl,', No Iterations ',rep%iters
end subroutine Diagonally_Preconditioned_CG

! `call fluxuvw(idew,idns,idtb, arvx,arvy,arvz, fif, acfe,acfw,fcf)` for one
! direction: the module scheme logicals and gds(iu) select the deferred
! correction; su, sv, sw (coef) gain the face terms; bp, bt, bb (coefb) are
! the east cell's scratch and zero on return.
subroutine fluxuvw(idew, idns, idtb, arvx, arvy, arvz, fif, acfe, acfw, fcf)
  use types; use parameters; use indexes; use coef; use coefb; use variables
  use iso_c_binding; use cap_hip
  implicit none
  integer, intent(in) :: idew, idns, idtb
  real(prec), dimension(nxyza) :: arvx, arvy, arvz, fif, acfe, acfw, fcf
  type(cap_fields) :: f
  integer(c_int) :: sch
  call cap_fields_of_modules(f)
  sch = 0
  if (lcds) sch = ior(sch, CAP_UVW_CDS)
  if (lcds4) sch = ior(sch, CAP_UVW_CDS4)
  if (lluds) sch = ior(sch, CAP_UVW_LUDS)
  if (lquds) sch = ior(sch, CAP_UVW_QUICK)
  if (flux_limiter) sch = ior(sch, CAP_UVW_LIMITER)
  if (lsmart) sch = ior(sch, CAP_UVW_SMART)
  if (lavl) sch = ior(sch, CAP_UVW_AVL)
  if (lmuscl) sch = ior(sch, CAP_UVW_MUSCL)
  if (lumist) sch = ior(sch, CAP_UVW_UMIST)
  if (lgamma) sch = ior(sch, CAP_UVW_GAMMA)
  call cap_check(cap_fluxuvw(cap_context(ni, nj, nk), f, cap_addr(vis), int(idew, c_int), &
                 int(idns, c_int), int(idtb, c_int), cap_addr(arvx), cap_addr(arvy), &
                 cap_addr(arvz), cap_addr(fif), cap_addr(acfe), cap_addr(acfw), cap_addr(fcf), &
                 cap_addr(su), cap_addr(sv), cap_addr(sw), cap_addr(bp), cap_addr(bt), &
                 cap_addr(bb), sch, real(gds(iu), c_double)), 'fluxuvw')
end subroutine fluxuvw

! `call fluxmc(inp,idew,idns,idtb,fif,fmcor)`: the reference computes one
! face per call, inside the cell loop of calcp-multiple_correction_SIMPLE.f90:
! 126-158.  The device computes the face of every cell of k,i,j = 3..n?mm at
! once: a call for the loop's first cell (lk(3)+li(3)+3) recomputes the
! direction's values, every other call reads them.  The inputs (nodes,
! gradients of p', fif) must therefore not change within one such loop, which
! holds in calcp, where the loop changes only f1..f3 and su.
subroutine fluxmc(inp, idew, idns, idtb, fif, fmcor)
  use types; use parameters; use indexes
  use iso_c_binding; use cap_hip
  implicit none
  integer, intent(in) :: inp, idew, idns, idtb
  real(prec), dimension(nxyza), intent(in) :: fif
  real(prec), intent(inout) :: fmcor
  real(c_double), allocatable, target, save :: mc(:, :)
  type(cap_fields) :: f
  integer :: d
  if (idew == nj) then
    d = 1
  else if (idew == 1) then
    d = 2
  else
    d = 3
  end if
  if (.not. allocated(mc)) allocate(mc(nxyza, 3))
  if (size(mc, 1) /= nxyza) then
    deallocate(mc)
    allocate(mc(nxyza, 3))
  end if
  if (inp == lk(3) + li(3) + 3) then
    call cap_fields_of_modules(f)
    call cap_check(cap_fluxmc(cap_context(ni, nj, nk), f, int(idew, c_int), int(idns, c_int), &
                              int(idtb, c_int), cap_caddr(fif), c_loc(mc(1, d))), 'fluxmc')
  end if
  fmcor = mc(inp, d)
end subroutine fluxmc

! `call update_values_at_ghostcells` (update_values_at_ghostcells_nonperiodic.f90:
! 3-642): ghost values of u, v, w, te, ed, t, vis by the boundary types of
! module bc and the inlet values of module inlet.
subroutine update_values_at_ghostcells
  use types; use parameters; use indexes; use geometry; use variables; use bc; use inlet
  use iso_c_binding; use cap_hip
  implicit none
  type(c_ptr) :: var(7), g(12)
  type(cap_bc) :: b
  var = [cap_addr(u), cap_addr(v), cap_addr(w), cap_addr(te), cap_addr(ed), cap_addr(t), &
         cap_addr(vis)]
  g = [cap_addr(fx), cap_addr(fy), cap_addr(fz), cap_addr(ar1x), cap_addr(ar1y), cap_addr(ar1z), &
       cap_addr(ar2x), cap_addr(ar2y), cap_addr(ar2z), cap_addr(ar3x), cap_addr(ar3y), &
       cap_addr(ar3z)]
  b%lbw = cap_iaddr(lbw); b%lbe = cap_iaddr(lbe); b%lbs = cap_iaddr(lbs)
  b%lbn = cap_iaddr(lbn); b%lbb = cap_iaddr(lbb); b%lbt = cap_iaddr(lbt)
  b%uin = uin; b%vin = vin; b%win = win; b%tein = tein; b%edin = edin; b%tin = tin
  call cap_check(cap_update_ghostcells(cap_context(ni, nj, nk), var, g, b), &
                 'update_values_at_ghostcells')
end subroutine update_values_at_ghostcells
