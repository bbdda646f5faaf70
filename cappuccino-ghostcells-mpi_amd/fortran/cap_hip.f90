!-----------------------------------------------------------------------
! cap_hip.f90 -- ISO_C_BINDING interface to the MI355X p' library
! (include/cap_amd.h).  Self-contained: needs only iso_c_binding.
!
! The C structs cap_coef (16 pointers) and cap_report map to the bind(c)
! derived types below; every entry point returns an integer status.
!-----------------------------------------------------------------------
module cap_hip
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: CAP_OK = 0, CAP_MAXITER = 1, CAP_NAN = 2, &
                               CAP_GUESS_REJECTED = 3
  integer(c_int), parameter :: CAP_ERR_ARG = -1, CAP_ERR_HIP = -2, CAP_ERR_TIMEOUT = -3, &
                               CAP_ERR_NOMEM = -4, CAP_ERR_NODEV = -5
  integer(c_int), parameter :: CAP_DEVICE_POINTERS = 1
  ! convection schemes of cap_fluxuvw (the reference's lcds, lcds4, lluds,
  ! lquds, flux_limiter, lsmart, lavl, lmuscl, lumist, lgamma)
  integer(c_int), parameter :: CAP_UVW_CDS = 1, CAP_UVW_CDS4 = 2, CAP_UVW_LUDS = 4, &
                               CAP_UVW_QUICK = 8, CAP_UVW_LIMITER = 16, CAP_UVW_SMART = 32, &
                               CAP_UVW_AVL = 64, CAP_UVW_MUSCL = 128, CAP_UVW_UMIST = 256, &
                               CAP_UVW_GAMMA = 512, CAP_UVW_FP32 = 1024

  ! module `coef` + `coefb` arrays (modules_allocatable.f90:161-180)
  type, bind(c) :: cap_coef
    type(c_ptr) :: ae = c_null_ptr, aw = c_null_ptr, an = c_null_ptr, as = c_null_ptr
    type(c_ptr) :: at = c_null_ptr, ab = c_null_ptr, ap = c_null_ptr, su = c_null_ptr
    type(c_ptr) :: res = c_null_ptr
    type(c_ptr) :: bb = c_null_ptr, bs = c_null_ptr, bw = c_null_ptr, bp = c_null_ptr
    type(c_ptr) :: bn = c_null_ptr, be = c_null_ptr, bt = c_null_ptr
  end type cap_coef

  ! fields read by the p' assembly (modules geometry, variables, coef, gradients)
  type, bind(c) :: cap_fields
    type(c_ptr) :: x = c_null_ptr, y = c_null_ptr, z = c_null_ptr
    type(c_ptr) :: xc = c_null_ptr, yc = c_null_ptr, zc = c_null_ptr, vol = c_null_ptr
    type(c_ptr) :: fx = c_null_ptr, fy = c_null_ptr, fz = c_null_ptr
    type(c_ptr) :: ar1x = c_null_ptr, ar1y = c_null_ptr, ar1z = c_null_ptr
    type(c_ptr) :: ar2x = c_null_ptr, ar2y = c_null_ptr, ar2z = c_null_ptr
    type(c_ptr) :: ar3x = c_null_ptr, ar3y = c_null_ptr, ar3z = c_null_ptr
    type(c_ptr) :: den = c_null_ptr, apu = c_null_ptr, apv = c_null_ptr, apw = c_null_ptr
    type(c_ptr) :: u = c_null_ptr, v = c_null_ptr, w = c_null_ptr, p = c_null_ptr
    type(c_ptr) :: gradu = c_null_ptr, gradv = c_null_ptr, gradw = c_null_ptr, gradp = c_null_ptr
  end type cap_fields

  ! boundary types per boundary cell and inlet values (modules bc, inlet)
  type, bind(c) :: cap_bc
    type(c_ptr) :: lbw = c_null_ptr, lbe = c_null_ptr, lbs = c_null_ptr, lbn = c_null_ptr
    type(c_ptr) :: lbb = c_null_ptr, lbt = c_null_ptr
    real(c_double) :: uin = 0, vin = 0, win = 0, tein = 0, edin = 0, tin = 0
  end type cap_bc

  ! module state of the pressure-correction drivers (device pointers)
  type, bind(c) :: cap_flow
    type(c_ptr) :: x = c_null_ptr, y = c_null_ptr, z = c_null_ptr
    type(c_ptr) :: xc = c_null_ptr, yc = c_null_ptr, zc = c_null_ptr, vol = c_null_ptr
    type(c_ptr) :: fx = c_null_ptr, fy = c_null_ptr, fz = c_null_ptr
    type(c_ptr) :: ar1x = c_null_ptr, ar1y = c_null_ptr, ar1z = c_null_ptr
    type(c_ptr) :: ar2x = c_null_ptr, ar2y = c_null_ptr, ar2z = c_null_ptr
    type(c_ptr) :: ar3x = c_null_ptr, ar3y = c_null_ptr, ar3z = c_null_ptr
    type(c_ptr) :: den = c_null_ptr
    type(c_ptr) :: u = c_null_ptr, v = c_null_ptr, w = c_null_ptr, p = c_null_ptr, pp = c_null_ptr
    type(c_ptr) :: f1 = c_null_ptr, f2 = c_null_ptr, f3 = c_null_ptr
    type(c_ptr) :: ae = c_null_ptr, aw = c_null_ptr, an = c_null_ptr, as = c_null_ptr
    type(c_ptr) :: at = c_null_ptr, ab = c_null_ptr, ap = c_null_ptr
    type(c_ptr) :: su = c_null_ptr, sv = c_null_ptr, sw = c_null_ptr, sp = c_null_ptr
    type(c_ptr) :: apu = c_null_ptr, apv = c_null_ptr, apw = c_null_ptr
    type(c_ptr) :: res = c_null_ptr, bb = c_null_ptr, bs = c_null_ptr, bw = c_null_ptr
    type(c_ptr) :: bp = c_null_ptr, bn = c_null_ptr, be = c_null_ptr, bt = c_null_ptr
    type(c_ptr) :: he = c_null_ptr, hw = c_null_ptr, hn = c_null_ptr, hs = c_null_ptr
    type(c_ptr) :: ht = c_null_ptr, hb = c_null_ptr
    type(c_ptr) :: gradu = c_null_ptr, gradv = c_null_ptr, gradw = c_null_ptr, gradp = c_null_ptr
    type(c_ptr) :: uo = c_null_ptr, vo = c_null_ptr, wo = c_null_ptr
    type(c_ptr) :: uoo = c_null_ptr, voo = c_null_ptr, woo = c_null_ptr
    type(c_ptr) :: t = c_null_ptr, te = c_null_ptr, ed = c_null_ptr, vis = c_null_ptr
    type(c_ptr) :: vart = c_null_ptr, con = c_null_ptr
    type(c_ptr) :: utt = c_null_ptr, vtt = c_null_ptr, wtt = c_null_ptr
  end type cap_flow

  ! the driver settings (modules parameters, boundc, time_mod)
  type, bind(c) :: cap_drv_opts
    integer(c_int) :: ncorr = 1, npcor = 1, solver = 5, nsw = 1000
    real(c_double) :: urf = 0.2d0, sor = 1d-8, alfa = 0.92d0
    integer(c_int) :: ijkpr = 1, const_mflux = 1
    integer(c_int) :: lout = 0, low = 0, loe = 0, los = 0, lon = 0, lob = 0, lot = 0
    real(c_double) :: flowin = 0
    type(c_ptr) :: lbw = c_null_ptr, lbe = c_null_ptr, lbs = c_null_ptr, lbn = c_null_ptr
    type(c_ptr) :: lbb = c_null_ptr, lbt = c_null_ptr
    integer(c_int) :: cn = 0, bdf = 0, buoy = 0, boussinesq = 0
    real(c_double) :: timestep = 0, btime = 0, densit = 1, beta = 0, tref = 0
    real(c_double) :: gravx = 0, gravy = 0, gravz = 0, cumulative = 0
  end type cap_drv_opts

  ! calcuvw settings beyond cap_drv_opts
  type, bind(c) :: cap_uvw_opts
    integer(c_int) :: schemes = 0
    real(c_double) :: gds = 1
    real(c_double) :: urfr(3) = 1, urfm(3) = 0, sor(3) = 0.1d0
    integer(c_int) :: nsw(3) = 5
    real(c_double) :: gradpcmf = 0
    integer(c_int) :: nx = 0, ny = 0, nz = 0
    integer(c_int) :: top_k = 0      ! 0: the plane lk(nimm), as the reference writes it
  end type cap_uvw_opts

  ! calc_vis_les settings
  type, bind(c) :: cap_les_opts
    real(c_double) :: viscos = 0, densit = 1, cappa = 0.41d0, elog = 8.432d0, urfvis = 1
    integer(c_int) :: nx = 0, ny = 0, nz = 0
    integer(c_int) :: top_k = 0
  end type cap_les_opts

  type, bind(c) :: cap_report
    real(c_double) :: resor, resl
    integer(c_int) :: iters, status
    real(c_double) :: ms
  end type cap_report

  interface
    integer(c_int) function cap_init(ctx, ni, nj, nk, flags) bind(c, name='cap_init')
      import :: c_ptr, c_int
      type(c_ptr), intent(out) :: ctx
      integer(c_int), value :: ni, nj, nk, flags
    end function
    integer(c_int) function cap_finalize(ctx) bind(c, name='cap_finalize')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function cap_sipsol(ctx, fi, c, alfa, sor, nsw, rep) bind(c, name='cap_sipsol')
      import :: c_ptr, c_int, c_double, cap_coef, cap_report
      type(c_ptr), value :: ctx, fi
      type(cap_coef), intent(in) :: c
      real(c_double), value :: alfa, sor
      integer(c_int), value :: nsw
      type(cap_report), intent(out) :: rep
    end function
    integer(c_int) function cap_pcgsip(ctx, fi, c, alfa, sor, nsw, rep) bind(c, name='cap_pcgsip')
      import :: c_ptr, c_int, c_double, cap_coef, cap_report
      type(c_ptr), value :: ctx, fi
      type(cap_coef), intent(in) :: c
      real(c_double), value :: alfa, sor
      integer(c_int), value :: nsw
      type(cap_report), intent(out) :: rep
    end function
    integer(c_int) function cap_cgstab(ctx, fi, c, sor, nsw, rep) bind(c, name='cap_cgstab')
      import :: c_ptr, c_int, c_double, cap_coef, cap_report
      type(c_ptr), value :: ctx, fi
      type(cap_coef), intent(in) :: c
      real(c_double), value :: sor
      integer(c_int), value :: nsw
      type(cap_report), intent(out) :: rep
    end function
    integer(c_int) function cap_pcg_jacobi(ctx, fi, c, sor, nsw, rep) bind(c, name='cap_pcg_jacobi')
      import :: c_ptr, c_int, c_double, cap_coef, cap_report
      type(c_ptr), value :: ctx, fi
      type(cap_coef), intent(in) :: c
      real(c_double), value :: sor
      integer(c_int), value :: nsw
      type(cap_report), intent(out) :: rep
    end function
    integer(c_int) function cap_iccg(ctx, fi, c, sor, nsw, rep) bind(c, name='cap_iccg')
      import :: c_ptr, c_int, c_double, cap_coef, cap_report
      type(c_ptr), value :: ctx, fi
      type(cap_coef), intent(in) :: c
      real(c_double), value :: sor
      integer(c_int), value :: nsw
      type(cap_report), intent(out) :: rep
    end function
    integer(c_int) function cap_cgstab_sip(ctx, fi, c, alfa, sor, nsw, rep) &
        bind(c, name='cap_cgstab_sip')
      import :: c_ptr, c_int, c_double, cap_coef, cap_report
      type(c_ptr), value :: ctx, fi
      type(cap_coef), intent(in) :: c
      real(c_double), value :: alfa, sor
      integer(c_int), value :: nsw
      type(cap_report), intent(out) :: rep
    end function
    integer(c_int) function cap_pcgsip_mixed(ctx, fi, c, alfa, sor, nsw, inner_rtol, rep) &
        bind(c, name='cap_pcgsip_mixed')
      import :: c_ptr, c_int, c_double, cap_coef, cap_report
      type(c_ptr), value :: ctx, fi
      type(cap_coef), intent(in) :: c
      real(c_double), value :: alfa, sor, inner_rtol
      integer(c_int), value :: nsw
      type(cap_report), intent(out) :: rep
    end function
    integer(c_int) function cap_pcgsip_fcg(ctx, fi, c, alfa, sor, nsw, rep) &
        bind(c, name='cap_pcgsip_fcg')
      import :: c_ptr, c_int, c_double, cap_coef, cap_report
      type(c_ptr), value :: ctx, fi
      type(cap_coef), intent(in) :: c
      real(c_double), value :: alfa, sor
      integer(c_int), value :: nsw
      type(cap_report), intent(out) :: rep
    end function
    integer(c_int) function cap_grad_gauss(ctx, u, geom, dudx, dudy, dudz) &
        bind(c, name='cap_grad_gauss')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, u, dudx, dudy, dudz
      type(c_ptr), intent(in) :: geom(13)
    end function
    integer(c_int) function cap_bpres(ctx, p, fx, fy, fz) bind(c, name='cap_bpres')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, p, fx, fy, fz
    end function
    integer(c_int) function cap_fluxmass(ctx, f, idew, idns, idtb, arvx, arvy, arvz, fif, &
        acfe, acfw, fcf) bind(c, name='cap_fluxmass')
      import :: c_ptr, c_int, cap_fields
      type(c_ptr), value :: ctx, arvx, arvy, arvz, fif, acfe, acfw, fcf
      type(cap_fields), intent(in) :: f
      integer(c_int), value :: idew, idns, idtb
    end function
    integer(c_int) function cap_piso_assemble_pressure_eq(ctx, f, idew, idns, idtb, arvx, arvy, &
        arvz, fif, acfe, acfw, fcf, ipc, bp, su) bind(c, name='cap_piso_assemble_pressure_eq')
      import :: c_ptr, c_int, cap_fields
      type(c_ptr), value :: ctx, arvx, arvy, arvz, fif, acfe, acfw, fcf, bp, su
      type(cap_fields), intent(in) :: f
      integer(c_int), value :: idew, idns, idtb, ipc
    end function
    integer(c_int) function cap_assemble_pressure_eq(ctx, f, idew, idns, idtb, arvx, arvy, arvz, &
        fif, acfe, acfw, fcf, su, sv, sw) bind(c, name='cap_assemble_pressure_eq')
      import :: c_ptr, c_int, cap_fields
      type(c_ptr), value :: ctx, arvx, arvy, arvz, fif, acfe, acfw, fcf, su, sv, sw
      type(cap_fields), intent(in) :: f
      integer(c_int), value :: idew, idns, idtb
    end function
    integer(c_int) function cap_fluxmc(ctx, f, idew, idns, idtb, fif, fmcor) bind(c, name='cap_fluxmc')
      import :: c_ptr, c_int, cap_fields
      type(c_ptr), value :: ctx, fif, fmcor
      type(cap_fields), intent(in) :: f
      integer(c_int), value :: idew, idns, idtb
    end function
    integer(c_int) function cap_fluxuvw(ctx, f, vis, idew, idns, idtb, arvx, arvy, arvz, fif, &
        acfe, acfw, fcf, su, sv, sw, bp, bt, bb, schemes, gds) bind(c, name='cap_fluxuvw')
      import :: c_ptr, c_int, c_double, cap_fields
      type(c_ptr), value :: ctx, vis, arvx, arvy, arvz, fif, acfe, acfw, fcf, su, sv, sw, bp, bt, bb
      type(cap_fields), intent(in) :: f
      integer(c_int), value :: idew, idns, idtb, schemes
      real(c_double), value :: gds
    end function
    ! the drivers' three per-direction calls in one pass (device pointers)
    integer(c_int) function cap_fluxmass3(ctx, f, ae, aw, an, as, at, ab, f1, f2, f3) &
        bind(c, name='cap_fluxmass3')
      import :: c_ptr, c_int, cap_fields
      type(c_ptr), value :: ctx, ae, aw, an, as, at, ab, f1, f2, f3
      type(cap_fields), intent(in) :: f
    end function
    integer(c_int) function cap_piso_assemble_pressure_eq3(ctx, f, ae, aw, an, as, at, ab, &
        f1, f2, f3, ipc, be, bn, bt, su) bind(c, name='cap_piso_assemble_pressure_eq3')
      import :: c_ptr, c_int, cap_fields
      type(c_ptr), value :: ctx, ae, aw, an, as, at, ab, f1, f2, f3, be, bn, bt, su
      type(cap_fields), intent(in) :: f
      integer(c_int), value :: ipc
    end function
    integer(c_int) function cap_fluxuvw3(ctx, f, vis, ae, aw, an, as, at, ab, f1, f2, f3, &
        su, sv, sw, bp, bt, bb, schemes, gds) bind(c, name='cap_fluxuvw3')
      import :: c_ptr, c_int, c_double, cap_fields
      type(c_ptr), value :: ctx, vis, ae, aw, an, as, at, ab, f1, f2, f3, su, sv, sw, bp, bt, bb
      type(cap_fields), intent(in) :: f
      integer(c_int), value :: schemes
      real(c_double), value :: gds
    end function
    integer(c_int) function cap_courant(ctx, f1, f2, f3, vol, su, timestep, time, conumfix, &
        conumfixvalue, conum) bind(c, name='cap_courant')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx, f1, f2, f3, vol, su
      real(c_double), intent(inout) :: timestep, time
      integer(c_int), value :: conumfix
      real(c_double), value :: conumfixvalue
      real(c_double), intent(out) :: conum(2)
    end function
    integer(c_int) function cap_courant_noalloc(ctx, f1, f2, f3, vol, su, timestep, time, &
        conumfix, conumfixvalue, conum) bind(c, name='cap_courant_noalloc')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx, f1, f2, f3, vol, su
      real(c_double), intent(inout) :: timestep, time
      integer(c_int), value :: conumfix
      real(c_double), value :: conumfixvalue
      real(c_double), intent(out) :: conum(2)
    end function
    integer(c_int) function cap_face_areas(ctx, x, y, z, ar) bind(c, name='cap_face_areas')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, x, y, z
      type(c_ptr), intent(in) :: ar(9)
    end function
    integer(c_int) function cap_intfac(ctx, nodes_centres, fx, fy, fz) bind(c, name='cap_intfac')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      type(c_ptr), intent(in) :: nodes_centres(6)
      type(c_ptr), value :: fx, fy, fz
    end function
    integer(c_int) function cap_grid_dims(path, ni, nj, nk) bind(c, name='cap_grid_dims')
      import :: c_char, c_int
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), intent(out) :: ni, nj, nk
    end function
    integer(c_int) function cap_read_grid(path, ni, nj, nk, arrays, lb, flags) &
        bind(c, name='cap_read_grid')
      import :: c_char, c_int, c_ptr
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: ni, nj, nk
      type(c_ptr), intent(in) :: arrays(19), lb(6)
      integer(c_int), intent(out) :: flags(9)
    end function
    integer(c_int) function cap_update_ghostcells_periodic(ctx, var, fz) &
        bind(c, name='cap_update_ghostcells_periodic')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, fz
      type(c_ptr), intent(in) :: var(7)
    end function
    integer(c_int) function cap_update_ghostcells(ctx, var, geom, bc) &
        bind(c, name='cap_update_ghostcells')
      import :: c_ptr, c_int, cap_bc
      type(c_ptr), value :: ctx
      type(c_ptr), intent(in) :: var(7), geom(12)
      type(cap_bc), intent(in) :: bc
    end function
    ! learned initial guess (no counterpart in the reference)
    integer(c_int) function cap_unet_create(ctx, nx, ny, nz, seed, net) bind(c, name='cap_unet_create')
      import :: c_ptr, c_int, c_long_long
      type(c_ptr), value :: ctx
      integer(c_int), value :: nx, ny, nz
      integer(c_long_long), value :: seed
      type(c_ptr), intent(out) :: net
    end function
    integer(c_int) function cap_unet_destroy(net) bind(c, name='cap_unet_destroy')
      import :: c_ptr, c_int
      type(c_ptr), value :: net
    end function
    integer(c_int) function cap_unet_num_params(net, count) bind(c, name='cap_unet_num_params')
      import :: c_ptr, c_int, c_long
      type(c_ptr), value :: net
      integer(c_long), intent(out) :: count
    end function
    integer(c_int) function cap_unet_get_params(net, params) bind(c, name='cap_unet_get_params')
      import :: c_ptr, c_int
      type(c_ptr), value :: net, params
    end function
    integer(c_int) function cap_unet_set_params(net, params) bind(c, name='cap_unet_set_params')
      import :: c_ptr, c_int
      type(c_ptr), value :: net, params
    end function
    integer(c_int) function cap_unet_forward(net, in, out) bind(c, name='cap_unet_forward')
      import :: c_ptr, c_int
      type(c_ptr), value :: net, in, out
    end function
    integer(c_int) function cap_unet_gradients(net, in, target, loss, grads) &
        bind(c, name='cap_unet_gradients')
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: net, in, target, grads
      real(c_float), intent(out) :: loss
    end function
    integer(c_int) function cap_unet_train_step(net, in, target, lr, loss) &
        bind(c, name='cap_unet_train_step')
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: net, in, target
      real(c_float), value :: lr
      real(c_float), intent(out) :: loss
    end function
    integer(c_int) function cap_unet_guess(ctx, net, fi, prev, c, scales) bind(c, name='cap_unet_guess')
      import :: c_ptr, c_int, c_double, cap_coef
      type(c_ptr), value :: ctx, net, fi, prev
      type(cap_coef), intent(in) :: c
      real(c_double), intent(out) :: scales(5)
    end function
    integer(c_int) function cap_pcgsip_guess(ctx, net, fi, prev, c, alfa, sor, nsw, rep, accepted) &
        bind(c, name='cap_pcgsip_guess')
      import :: c_ptr, c_int, c_double, cap_coef, cap_report
      type(c_ptr), value :: ctx, net, fi, prev
      type(cap_coef), intent(in) :: c
      real(c_double), value :: alfa, sor
      integer(c_int), value :: nsw
      type(cap_report), intent(out) :: rep
      integer(c_int), intent(out) :: accepted
    end function
    integer(c_int) function cap_continuity(ctx, f1, f2, f3, sums) bind(c, name='cap_continuity')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx, f1, f2, f3
      real(c_double), intent(out) :: sums(2)
    end function
    integer(c_int) function cap_calcp(ctx, f, o, reps, cont) bind(c, name='cap_calcp')
      import :: c_ptr, c_int, c_double, cap_flow, cap_drv_opts, cap_report
      type(c_ptr), value :: ctx
      type(cap_flow), intent(in) :: f
      type(cap_drv_opts), intent(in) :: o
      type(cap_report), intent(out) :: reps(*)
      real(c_double), intent(out) :: cont(*)
    end function
    integer(c_int) function cap_piso(ctx, f, o, pimple, reps, cont) bind(c, name='cap_piso')
      import :: c_ptr, c_int, c_double, cap_flow, cap_drv_opts, cap_report
      type(c_ptr), value :: ctx
      type(cap_flow), intent(in) :: f
      type(cap_drv_opts), intent(in) :: o
      integer(c_int), value :: pimple
      type(cap_report), intent(out) :: reps(*)
      real(c_double), intent(out) :: cont(*)
    end function
    integer(c_int) function cap_piso_guess(ctx, f, o, pimple, net, prev, reps, cont, accepted, l1) &
        bind(c, name='cap_piso_guess')
      import :: c_ptr, c_int, c_double, cap_flow, cap_drv_opts, cap_report
      type(c_ptr), value :: ctx, net, prev
      type(cap_flow), intent(in) :: f
      type(cap_drv_opts), intent(in) :: o
      integer(c_int), value :: pimple
      type(cap_report), intent(out) :: reps(*)
      real(c_double), intent(out) :: cont(*)
      integer(c_int), intent(out) :: accepted
      real(c_double), intent(out) :: l1(2)
    end function
    integer(c_int) function cap_outbc(ctx, f, o) bind(c, name='cap_outbc')
      import :: c_ptr, c_int, cap_flow, cap_drv_opts
      type(c_ptr), value :: ctx
      type(cap_flow), intent(in) :: f
      type(cap_drv_opts), intent(in) :: o
    end function
    integer(c_int) function cap_correct_bc(ctx, f) bind(c, name='cap_correct_bc')
      import :: c_ptr, c_int, cap_flow
      type(c_ptr), value :: ctx
      type(cap_flow), intent(in) :: f
    end function
    integer(c_int) function cap_calcuvw(ctx, f, spv, o, q, reps) bind(c, name='cap_calcuvw')
      import :: c_ptr, c_int, cap_flow, cap_drv_opts, cap_uvw_opts, cap_report
      type(c_ptr), value :: ctx, spv
      type(cap_flow), intent(in) :: f
      type(cap_drv_opts), intent(in) :: o
      type(cap_uvw_opts), intent(in) :: q
      type(cap_report), intent(out) :: reps(3)
    end function
    integer(c_int) function cap_find_strain_rate(ctx, f, strain, vorticity) &
        bind(c, name='cap_find_strain_rate')
      import :: c_ptr, c_int, cap_flow
      type(c_ptr), value :: ctx, strain, vorticity
      type(cap_flow), intent(in) :: f
    end function
    integer(c_int) function cap_calc_vis_les(ctx, f, strain, walldist, o) &
        bind(c, name='cap_calc_vis_les')
      import :: c_ptr, c_int, cap_flow, cap_les_opts
      type(c_ptr), value :: ctx, strain, walldist
      type(cap_flow), intent(in) :: f
      type(cap_les_opts), intent(in) :: o
    end function
    integer(c_int) function cap_gethbya(ctx, f, o) bind(c, name='cap_gethbya')
      import :: c_ptr, c_int, cap_flow, cap_drv_opts
      type(c_ptr), value :: ctx
      type(cap_flow), intent(in) :: f
      type(cap_drv_opts), intent(in) :: o
    end function
  end interface

  ! one context per process, created on first use for the current extents
  type(c_ptr), save :: cap_ctx_handle = c_null_ptr
  integer, save :: cap_ctx_dims(3) = 0

contains

  ! the library context for arrays of extents ni x nj x nk (host pointers)
  function cap_context(ni, nj, nk) result(ctx)
    integer, intent(in) :: ni, nj, nk
    type(c_ptr) :: ctx
    integer(c_int) :: rc
    if (c_associated(cap_ctx_handle)) then
      if (all(cap_ctx_dims == [ni, nj, nk])) then
        ctx = cap_ctx_handle
        return
      end if
      rc = cap_finalize(cap_ctx_handle)
      cap_ctx_handle = c_null_ptr
    end if
    rc = cap_init(cap_ctx_handle, int(ni, c_int), int(nj, c_int), int(nk, c_int), 0_c_int)
    if (rc /= CAP_OK) then
      write(*, '(a,i0)') 'cap_hip: cap_init failed with status ', rc
      stop 3
    end if
    cap_ctx_dims = [ni, nj, nk]
    ctx = cap_ctx_handle
  end function cap_context

  ! Pack the module arrays of the caller (explicit-shape dummies, so module
  ! allocatables without the TARGET attribute can be passed) into cap_coef.
  subroutine cap_pack(c, ae, aw, an, as, at, ab, ap, su, res, bb, bs, bw, bp, bn, be, bt)
    type(cap_coef), intent(out) :: c
    real(c_double), target, intent(inout) :: ae(*), aw(*), an(*), as(*), at(*), ab(*), ap(*), &
                                             su(*), res(*), bb(*), bs(*), bw(*), bp(*), &
                                             bn(*), be(*), bt(*)
    c%ae = c_loc(ae); c%aw = c_loc(aw); c%an = c_loc(an); c%as = c_loc(as)
    c%at = c_loc(at); c%ab = c_loc(ab); c%ap = c_loc(ap); c%su = c_loc(su)
    c%res = c_loc(res)
    c%bb = c_loc(bb); c%bs = c_loc(bs); c%bw = c_loc(bw); c%bp = c_loc(bp)
    c%bn = c_loc(bn); c%be = c_loc(be); c%bt = c_loc(bt)
  end subroutine cap_pack

  function cap_addr(x) result(p)
    real(c_double), target, intent(inout) :: x(*)
    type(c_ptr) :: p
    p = c_loc(x)
  end function cap_addr

  function cap_caddr(x) result(p)
    real(c_double), target, intent(in) :: x(*)
    type(c_ptr) :: p
    p = c_loc(x)
  end function cap_caddr

  function cap_iaddr(x) result(p)
    integer(c_int), target, intent(in) :: x(*)
    type(c_ptr) :: p
    p = c_loc(x)
  end function cap_iaddr

end module cap_hip
