"""Host-side binding of the MI355X p' library (``libcapamd.so``).

Mirrors the reference's solver interface: the reference calls
``call pcgsip(fi, ifi)`` with the matrix in module arrays ``ae..ap, su, res``
(module ``coef``) and ``bb..bt`` (module ``coefb``) -- see
modules_allocatable.f90:161-180 and pcg-sip.f90:3.  Here the module arrays
are an explicit :class:`Coef` object and the field id is replaced by the
explicit ``alfa / sor / nsw`` the reference reads from module ``parameters``.

Arrays are float64 in the reference's padded j-fastest layout (length
``ni*nj*nk``): numpy arrays for a host-pointer context, or torch tensors on
``cuda`` for a device-pointer context.  The library must be built
(``__graft_entry__.build()``); a missing library raises at import time --
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcapamd.so")

CAP_OK, CAP_MAXITER, CAP_NAN, CAP_GUESS_REJECTED = 0, 1, 2, 3
CAP_ERR_ARG, CAP_ERR_HIP, CAP_ERR_TIMEOUT, CAP_ERR_NOMEM, CAP_ERR_NODEV = -1, -2, -3, -4, -5
CAP_DEVICE_POINTERS = 1

COEF_FIELDS = ("ae", "aw", "an", "as_", "at", "ab", "ap", "su", "res",
               "bb", "bs", "bw", "bp", "bn", "be", "bt")
GEOM_FIELDS = ("fx", "fy", "fz", "ar1x", "ar1y", "ar1z", "ar2x", "ar2y", "ar2z",
               "ar3x", "ar3y", "ar3z", "vol")

# every entry point declared in include/cap_amd.h
EXPORTS = ("cap_init", "cap_finalize", "cap_last_error", "cap_sipsol", "cap_pcgsip",
           "cap_iccg", "cap_cgstab_sip", "cap_pcgsip_mixed", "cap_grad_gauss", "cap_bpres",
           "cap_build_lognormal", "cap_time_kernels", "cap_fluxmass",
           "cap_piso_assemble_pressure_eq", "cap_assemble_pressure_eq", "cap_fluxmc",
           "cap_continuity", "cap_fluxuvw", "cap_update_ghostcells",
           "cap_courant", "cap_unet_create", "cap_unet_destroy", "cap_unet_num_params",
           "cap_unet_get_params", "cap_unet_set_params", "cap_unet_forward",
           "cap_unet_gradients", "cap_unet_train_step", "cap_unet_guess", "cap_pcgsip_guess",
           "cap_calcp", "cap_piso", "cap_outbc", "cap_correct_bc", "cap_gethbya",
           "cap_calcuvw", "cap_find_strain_rate", "cap_calc_vis_les", "cap_courant_noalloc",
           "cap_face_areas", "cap_intfac", "cap_grid_dims", "cap_read_grid", "cap_cgstab",
           "cap_pcg_jacobi", "cap_update_ghostcells_periodic", "cap_piso_guess",
           "cap_pcgsip_fcg", "cap_fluxmass3", "cap_piso_assemble_pressure_eq3",
           "cap_fluxuvw3")
FIELD_NAMES = ("x", "y", "z", "xc", "yc", "zc", "vol", "fx", "fy", "fz", "ar1x", "ar1y",
               "ar1z", "ar2x", "ar2y", "ar2z", "ar3x", "ar3y", "ar3z", "den", "apu", "apv",
               "apw", "u", "v", "w", "p", "gradu", "gradv", "gradw", "gradp")


class CapError(RuntimeError):
    pass


class _Coef(C.Structure):
    _fields_ = [(f, C.c_void_p) for f in COEF_FIELDS]


class _Fields(C.Structure):
    _fields_ = [(f, C.c_void_p) for f in FIELD_NAMES]


class _Bc(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("lbw", "lbe", "lbs", "lbn", "lbb", "lbt")] + \
               [(n, C.c_double) for n in ("uin", "vin", "win", "tein", "edin", "tin")]


# module state of the pressure-correction drivers, in the order of cap_flow
FLOW_NAMES = ("x", "y", "z", "xc", "yc", "zc", "vol", "fx", "fy", "fz",
              "ar1x", "ar1y", "ar1z", "ar2x", "ar2y", "ar2z", "ar3x", "ar3y", "ar3z",
              "den", "u", "v", "w", "p", "pp", "f1", "f2", "f3",
              "ae", "aw", "an", "as", "at", "ab", "ap", "su", "sv", "sw", "sp", "apu", "apv", "apw",
              "res", "bb", "bs", "bw", "bp", "bn", "be", "bt",
              "he", "hw", "hn", "hs", "ht", "hb",
              "gradu", "gradv", "gradw", "gradp",
              "uo", "vo", "wo", "uoo", "voo", "woo",
              "t", "te", "ed", "vis", "vart", "con", "utt", "vtt", "wtt")


class _Flow(C.Structure):
    _fields_ = [(f, C.c_void_p) for f in FLOW_NAMES]


class DrvOpts(C.Structure):
    """Settings of the drivers (cap_drv_opts): module parameters, boundc and
    time_mod values the reference routines read."""
    _fields_ = [("ncorr", C.c_int), ("npcor", C.c_int), ("solver", C.c_int), ("nsw", C.c_int),
                ("urf", C.c_double), ("sor", C.c_double), ("alfa", C.c_double),
                ("ijkpr", C.c_int), ("const_mflux", C.c_int),
                ("lout", C.c_int), ("low", C.c_int), ("loe", C.c_int), ("los", C.c_int),
                ("lon", C.c_int), ("lob", C.c_int), ("lot", C.c_int),
                ("flowin", C.c_double)] + \
               [(n, C.c_void_p) for n in ("lbw", "lbe", "lbs", "lbn", "lbb", "lbt")] + \
               [("cn", C.c_int), ("bdf", C.c_int), ("buoy", C.c_int), ("boussinesq", C.c_int)] + \
               [(n, C.c_double) for n in ("timestep", "btime", "densit", "beta", "tref",
                                          "gravx", "gravy", "gravz", "cumulative")]


class UvwOpts(C.Structure):
    """calcuvw's settings beyond DrvOpts (cap_uvw_opts)."""
    _fields_ = [("schemes", C.c_int), ("gds", C.c_double), ("urfr", C.c_double * 3),
                ("urfm", C.c_double * 3), ("sor", C.c_double * 3), ("nsw", C.c_int * 3),
                ("gradpcmf", C.c_double), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("top_k", C.c_int)]


class LesOpts(C.Structure):
    """calc_vis_les settings (cap_les_opts)."""
    _fields_ = [("viscos", C.c_double), ("densit", C.c_double), ("cappa", C.c_double),
                ("elog", C.c_double), ("urfvis", C.c_double), ("nx", C.c_int), ("ny", C.c_int),
                ("nz", C.c_int), ("top_k", C.c_int)]


class _Report(C.Structure):
    _fields_ = [("resor", C.c_double), ("resl", C.c_double), ("iters", C.c_int),
                ("status", C.c_int), ("ms", C.c_double)]


_lib = None


def load():
    """Load libcapamd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CapError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    L = C.CDLL(LIB_PATH)
    vp, i, d = C.c_void_p, C.c_int, C.c_double
    L.cap_init.argtypes = [C.POINTER(vp), i, i, i, i]
    L.cap_finalize.argtypes = [vp]
    L.cap_last_error.argtypes = [vp]
    L.cap_last_error.restype = C.c_char_p
    for fn in ("cap_sipsol", "cap_pcgsip", "cap_cgstab_sip"):
        getattr(L, fn).argtypes = [vp, vp, C.POINTER(_Coef), d, d, i, C.POINTER(_Report)]
    L.cap_iccg.argtypes = [vp, vp, C.POINTER(_Coef), d, i, C.POINTER(_Report)]
    L.cap_cgstab.argtypes = [vp, vp, C.POINTER(_Coef), d, i, C.POINTER(_Report)]
    L.cap_pcg_jacobi.argtypes = [vp, vp, C.POINTER(_Coef), d, i, C.POINTER(_Report)]
    L.cap_pcgsip_mixed.argtypes = [vp, vp, C.POINTER(_Coef), d, d, i, d, C.POINTER(_Report)]
    L.cap_pcgsip_fcg.argtypes = [vp, vp, C.POINTER(_Coef), d, d, i, C.POINTER(_Report)]
    L.cap_grad_gauss.argtypes = [vp, vp, C.POINTER(vp), vp, vp, vp]
    L.cap_bpres.argtypes = [vp, vp, vp, vp, vp]
    L.cap_build_lognormal.argtypes = [vp, C.c_ulonglong, d, C.POINTER(_Coef)]
    L.cap_time_kernels.argtypes = [vp, i, i, C.POINTER(d)]
    F = C.POINTER(_Fields)
    L.cap_fluxmass.argtypes = [vp, F, i, i, i, vp, vp, vp, vp, vp, vp, vp]
    L.cap_piso_assemble_pressure_eq.argtypes = [vp, F, i, i, i, vp, vp, vp, vp, vp, vp, vp, i,
                                                vp, vp]
    L.cap_assemble_pressure_eq.argtypes = [vp, F, i, i, i, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.cap_fluxmc.argtypes = [vp, F, i, i, i, vp, vp]
    L.cap_continuity.argtypes = [vp, vp, vp, vp, C.POINTER(d)]
    L.cap_fluxuvw.argtypes = [vp, F, vp, i, i, i] + [vp] * 13 + [i, d]
    L.cap_fluxmass3.argtypes = [vp, F] + [vp] * 9
    L.cap_piso_assemble_pressure_eq3.argtypes = [vp, F] + [vp] * 9 + [i] + [vp] * 4
    L.cap_fluxuvw3.argtypes = [vp, F] + [vp] * 16 + [i, d]
    L.cap_update_ghostcells.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(_Bc)]
    L.cap_update_ghostcells_periodic.argtypes = [vp, C.POINTER(vp), vp]
    L.cap_courant.argtypes = [vp, vp, vp, vp, vp, vp, C.POINTER(d), C.POINTER(d), i, d,
                              C.POINTER(d)]
    L.cap_courant_noalloc.argtypes = L.cap_courant.argtypes
    L.cap_face_areas.argtypes = [vp, vp, vp, vp, C.POINTER(vp)]
    L.cap_intfac.argtypes = [vp, C.POINTER(vp), vp, vp, vp]
    L.cap_grid_dims.argtypes = [C.c_char_p, C.POINTER(i), C.POINTER(i), C.POINTER(i)]
    L.cap_read_grid.argtypes = [C.c_char_p, i, i, i, C.POINTER(vp), C.POINTER(vp), C.POINTER(i)]
    f = C.c_float
    L.cap_unet_create.argtypes = [vp, i, i, i, C.c_ulonglong, C.POINTER(vp)]
    L.cap_unet_destroy.argtypes = [vp]
    L.cap_unet_num_params.argtypes = [vp, C.POINTER(C.c_long)]
    L.cap_unet_get_params.argtypes = [vp, vp]
    L.cap_unet_set_params.argtypes = [vp, vp]
    L.cap_unet_forward.argtypes = [vp, vp, vp]
    L.cap_unet_gradients.argtypes = [vp, vp, vp, C.POINTER(f), vp]
    L.cap_unet_train_step.argtypes = [vp, vp, vp, f, C.POINTER(f)]
    Fl, O = C.POINTER(_Flow), C.POINTER(DrvOpts)
    L.cap_calcp.argtypes = [vp, Fl, O, C.POINTER(_Report), C.POINTER(d)]
    L.cap_piso.argtypes = [vp, Fl, O, i, C.POINTER(_Report), C.POINTER(d)]
    L.cap_outbc.argtypes = [vp, Fl, O]
    L.cap_correct_bc.argtypes = [vp, Fl]
    L.cap_gethbya.argtypes = [vp, Fl, O]
    L.cap_calcuvw.argtypes = [vp, Fl, vp, O, C.POINTER(UvwOpts), C.POINTER(_Report)]
    L.cap_find_strain_rate.argtypes = [vp, Fl, vp, vp]
    L.cap_calc_vis_les.argtypes = [vp, Fl, vp, vp, C.POINTER(LesOpts)]
    L.cap_unet_guess.argtypes = [vp, vp, vp, vp, C.POINTER(_Coef), C.POINTER(d)]
    L.cap_pcgsip_guess.argtypes = [vp, vp, vp, vp, C.POINTER(_Coef), d, d, i, C.POINTER(_Report),
                                   C.POINTER(i)]
    L.cap_piso_guess.argtypes = [vp, Fl, O, i, vp, vp, C.POINTER(_Report), C.POINTER(d),
                                 C.POINTER(i), C.POINTER(d)]
    _lib = L
    return L


def _check_array(a, kind, n=None, device=None, what="array"):
    """Raise CapError unless `a` is a contiguous array of dtype `kind`
    ("float64", "float32" or "int32") with exactly n elements (when n is
    given) that lives on the device when `device` is True and on the host
    when it is False.  The C side reads and writes exactly n elements at the
    address, so every mismatch is refused here rather than left to overrun a
    buffer or to hand a host address to a kernel."""
    if hasattr(a, "data_ptr"):            # torch tensor
        import torch
        want = {"float64": torch.float64, "float32": torch.float32, "int32": torch.int32}[kind]
        if a.dtype != want:
            raise CapError(f"{what}: dtype {a.dtype}, expected {kind}")
        if not a.is_contiguous():
            raise CapError(f"{what}: not contiguous")
        if device is not None and bool(a.is_cuda) != bool(device):
            raise CapError(f"{what}: a {'device' if a.is_cuda else 'host'} tensor passed to a "
                           f"{'device' if device else 'host'}-pointer context")
        if n is not None and a.numel() != n:
            raise CapError(f"{what}: {a.numel()} elements, expected {n}")
        return a.data_ptr()
    import numpy as np
    if not isinstance(a, np.ndarray):
        raise CapError(f"{what}: {type(a).__name__} is neither a numpy array nor a torch tensor")
    if a.dtype != np.dtype(kind):
        raise CapError(f"{what}: dtype {a.dtype}, expected {kind}")
    if not a.flags.c_contiguous:
        raise CapError(f"{what}: not contiguous")
    if device:
        raise CapError(f"{what}: a host array passed to a device-pointer context")
    if n is not None and a.size != n:
        raise CapError(f"{what}: {a.size} elements, expected {n}")
    return a.ctypes.data


def _ptr(a, n=None, device=None, what="array"):
    """float64 array (numpy, or torch on the device); None stays NULL."""
    if a is None:
        return None
    return _check_array(a, "float64", n, device, what)


def _iptr(a, n=None, device=None, what="array"):
    """int32 array (numpy, or torch on the device)"""
    return _check_array(a, "int32", n, device, what)


class Coef:
    """The module arrays of `coef` and `coefb` that the solvers read/write."""

    def __init__(self, **arrays):
        for f in COEF_FIELDS:
            setattr(self, f, arrays.get(f))

    def _struct(self, ctx):
        return _Coef(*[ctx._f(getattr(self, f), f) for f in COEF_FIELDS])


class Context:
    """One solver context (one HIP stream) for arrays of extents ni x nj x nk."""

    def __init__(self, ni, nj, nk, device_pointers=False):
        self._lib = load()
        self.dims = (int(ni), int(nj), int(nk))
        self.device_pointers = bool(device_pointers)
        h = C.c_void_p()
        rc = self._lib.cap_init(C.byref(h), *self.dims,
                                CAP_DEVICE_POINTERS if device_pointers else 0)
        if rc != CAP_OK:
            raise CapError(f"cap_init failed with status {rc}")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.cap_finalize(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def n(self):
        ni, nj, nk = self.dims
        return ni * nj * nk

    def _f(self, a, what="array", mult=1):
        """Pointer of a float64 field of mult * ni*nj*nk elements on the
        context's side (device or host); None stays NULL."""
        return _ptr(a, mult * self.n, self.device_pointers, what)

    def _lb_sizes(self):
        ni, nj, nk = self.dims
        return (nj * nk, nj * nk, ni * nk, ni * nk, ni * nj, ni * nj)

    def error(self):
        return self._lib.cap_last_error(self._h).decode()

    def _check(self, rc, what, allow=(CAP_OK, CAP_MAXITER)):
        if rc not in allow:
            raise CapError(f"{what} failed with status {rc}: {self.error()}")
        return rc

    def _solve(self, fn, fi, coef, *args):
        rep = _Report()
        st = coef._struct(self)
        rc = fn(self._h, self._f(fi), C.byref(st), *args, C.byref(rep))
        self._check(rc, fn.__name__, allow=(CAP_OK, CAP_MAXITER, CAP_NAN))
        return dict(resor=rep.resor, resl=rep.resl, iters=rep.iters, status=rep.status,
                    converged=rep.status == CAP_OK, ms=rep.ms)

    # --- drop-ins for the reference solvers -------------------------------
    def sipsol(self, fi, coef, alfa=0.92, sor=1e-8, nsw=1000):
        return self._solve(self._lib.cap_sipsol, fi, coef, alfa, sor, nsw)

    def pcgsip(self, fi, coef, alfa=0.92, sor=1e-8, nsw=1000):
        return self._solve(self._lib.cap_pcgsip, fi, coef, alfa, sor, nsw)

    def iccg(self, fi, coef, sor=1e-8, nsw=1000):
        return self._solve(self._lib.cap_iccg, fi, coef, sor, nsw)

    def cgstab_sip(self, fi, coef, alfa=0.92, sor=1e-8, nsw=1000):
        return self._solve(self._lib.cap_cgstab_sip, fi, coef, alfa, sor, nsw)

    def cgstab(self, fi, coef, sor=1e-8, nsw=1000):
        """cgstab.f90: BiCGStab with the diagonal ILU(0) preconditioner"""
        return self._solve(self._lib.cap_cgstab, fi, coef, sor, nsw)

    def pcg_jacobi(self, fi, coef, sor=1e-8, nsw=1000):
        """pcg-jacobi.f90 (Diagonally_Preconditioned_CG)"""
        return self._solve(self._lib.cap_pcg_jacobi, fi, coef, sor, nsw)

    def pcgsip_mixed(self, fi, coef, alfa=0.92, sor=1e-8, nsw=5000, inner_rtol=1e-3):
        return self._solve(self._lib.cap_pcgsip_mixed, fi, coef, alfa, sor, nsw, inner_rtol)

    def pcgsip_fcg(self, fi, coef, alfa=0.92, sor=1e-8, nsw=1000):
        """PCG-SIP with the preconditioner in FP32 (flexible beta), FP64
        otherwise (cap_pcgsip_fcg)"""
        return self._solve(self._lib.cap_pcgsip_fcg, fi, coef, alfa, sor, nsw)

    # --- stencils -----------------------------------------------------------
    def grad_gauss(self, u, geom, dudx, dudy, dudz):
        arr = (C.c_void_p * 13)(*[self._f(geom[f]) for f in GEOM_FIELDS])
        rc = self._lib.cap_grad_gauss(self._h, self._f(u), arr, self._f(dudx), self._f(dudy), self._f(dudz))
        self._check(rc, "cap_grad_gauss", allow=(CAP_OK,))

    def bpres(self, p, fx, fy, fz):
        rc = self._lib.cap_bpres(self._h, self._f(p), self._f(fx), self._f(fy), self._f(fz))
        self._check(rc, "cap_bpres", allow=(CAP_OK,))

    def time_kernels(self, fp32=False, reps=20):
        """Average ms of the 6 PCG-SIP kernels (see cap_amd.h)."""
        ms = (C.c_double * 6)()
        rc = self._lib.cap_time_kernels(self._h, 1 if fp32 else 0, reps, ms)
        self._check(rc, "cap_time_kernels", allow=(CAP_OK,))
        return list(ms)

    # --- p' assembly (fields: dict of the FIELD_NAMES arrays) --------------
    def _dir(self, d):
        ni, nj, nk = self.dims
        return [(nj, 1, ni * nj), (1, ni * nj, nj), (ni * nj, nj, 1)][d]

    def _fields(self, f):
        return _Fields(*[self._f(f.get(q), q, 3 if q.startswith("grad") else 1)
                         for q in FIELD_NAMES])

    def fluxmass(self, f, d, acfe, acfw, fcf):
        """Direction d: 0 east (ar1, fx), 1 north (ar2, fy), 2 top (ar3, fz)."""
        arv = [f[f"ar{d + 1}{c}"] for c in "xyz"]
        st = self._fields(f)
        rc = self._lib.cap_fluxmass(self._h, C.byref(st), *self._dir(d), *map(self._f, arv),
                                    self._f(f[("fx", "fy", "fz")[d]]),
                                    self._f(acfe), self._f(acfw), self._f(fcf))
        self._check(rc, "cap_fluxmass", allow=(CAP_OK,))

    def piso_assemble_pressure_eq(self, f, d, acfe, acfw, fcf, ipc, bp, su):
        arv = [f[f"ar{d + 1}{c}"] for c in "xyz"]
        st = self._fields(f)
        rc = self._lib.cap_piso_assemble_pressure_eq(
            self._h, C.byref(st), *self._dir(d), *map(self._f, arv), self._f(f[("fx", "fy", "fz")[d]]),
            self._f(acfe), self._f(acfw), self._f(fcf), ipc, self._f(bp), self._f(su))
        self._check(rc, "cap_piso_assemble_pressure_eq", allow=(CAP_OK,))

    def assemble_pressure_eq(self, f, d, acfe, acfw, fcf, su, sv, sw):
        arv = [f[f"ar{d + 1}{c}"] for c in "xyz"]
        st = self._fields(f)
        rc = self._lib.cap_assemble_pressure_eq(
            self._h, C.byref(st), *self._dir(d), *map(self._f, arv), self._f(f[("fx", "fy", "fz")[d]]),
            self._f(acfe), self._f(acfw), self._f(fcf), self._f(su), self._f(sv), self._f(sw))
        self._check(rc, "cap_assemble_pressure_eq", allow=(CAP_OK,))

    def fluxuvw(self, f, d, vis, acfe, acfw, fcf, su, sv, sw, bp, bt, bb, schemes, gds):
        """Momentum face fluxes of direction d (0 east, 1 north, 2 top);
        schemes: CAP_UVW_* bits (the reference's lcds, lcds4, ... flags)."""
        arv = [f[f"ar{d + 1}{c}"] for c in "xyz"]
        st = self._fields(f)
        rc = self._lib.cap_fluxuvw(
            self._h, C.byref(st), self._f(vis), *self._dir(d), *map(self._f, arv),
            self._f(f[("fx", "fy", "fz")[d]]), self._f(acfe), self._f(acfw), self._f(fcf), self._f(su),
            self._f(sv), self._f(sw), self._f(bp), self._f(bt), self._f(bb), int(schemes), float(gds))
        self._check(rc, "cap_fluxuvw", allow=(CAP_OK,))

    def fluxmass3(self, f, a6, f3):
        """The three directions of fluxmass in one pass: a6 = [ae, aw, an, as,
        at, ab], f3 = [f1, f2, f3] (device arrays)."""
        st = self._fields(f)
        rc = self._lib.cap_fluxmass3(self._h, C.byref(st), *map(self._f, a6), *map(self._f, f3))
        self._check(rc, "cap_fluxmass3", allow=(CAP_OK,))

    def piso_assemble_pressure_eq3(self, f, a6, f3, ipc, bp3, su):
        """The three directions of piso_assemble_pressure_eq in one pass; bp3 =
        [be, bn, bt]."""
        st = self._fields(f)
        rc = self._lib.cap_piso_assemble_pressure_eq3(
            self._h, C.byref(st), *map(self._f, a6), *map(self._f, f3), int(ipc),
            *map(self._f, bp3), self._f(su))
        self._check(rc, "cap_piso_assemble_pressure_eq3", allow=(CAP_OK,))

    def fluxuvw3(self, f, vis, a6, f3, s3, scr3, schemes, gds):
        """calcuvw's three fluxuvw calls (CDS) in one pass: s3 = [su, sv, sw],
        scr3 = the scratch [bp, bt, bb]."""
        st = self._fields(f)
        rc = self._lib.cap_fluxuvw3(self._h, C.byref(st), self._f(vis), *map(self._f, a6),
                                    *map(self._f, f3), *map(self._f, s3), *map(self._f, scr3),
                                    int(schemes), float(gds))
        self._check(rc, "cap_fluxuvw3", allow=(CAP_OK,))

    def update_ghostcells(self, var, geom, lb, inl):
        """var: [u, v, w, te, ed, t, vis] (updated in place); geom: [fx, fy, fz,
        ar1x .. ar3z]; lb: int32 arrays [lbw, lbe, lbs, lbn, lbb, lbt]; inl:
        (uin, vin, win, tein, edin, tin)."""
        V = (C.c_void_p * 7)(*[self._f(a) for a in var])
        G = (C.c_void_p * 12)(*[self._f(a) for a in geom])
        bc = _Bc(*[_iptr(a, n, self.device_pointers, "lb")
                   for a, n in zip(lb, self._lb_sizes())], *[float(x) for x in inl])
        rc = self._lib.cap_update_ghostcells(self._h, V, G, C.byref(bc))
        self._check(rc, "cap_update_ghostcells", allow=(CAP_OK,))

    def update_ghostcells_periodic(self, var, fz):
        """var: [u, v, w, te, ed, t, vis] (t untouched); the periodic variant
        update_values_at_ghostcells.f90 (periodic i and j, walls at k)."""
        V = (C.c_void_p * 7)(*[self._f(a, q) for a, q in
                                zip(var, ("u", "v", "w", "te", "ed", "t", "vis"))])
        rc = self._lib.cap_update_ghostcells_periodic(self._h, V, self._f(fz, "fz"))
        self._check(rc, "cap_update_ghostcells_periodic", allow=(CAP_OK,))

    def courant(self, f1, f2, f3, vol, su, timestep, time, conumfix=False, conumfixvalue=0.5,
                noalloc=False):
        """Returns (CoNum, meanCoNum, timestep, time) after CourantNo.h.
        noalloc: through cap_courant_noalloc (static scratch, no device
        allocation; device-pointer contexts only)."""
        ts, tm = C.c_double(timestep), C.c_double(time)
        out = (C.c_double * 2)()
        fn = self._lib.cap_courant_noalloc if noalloc else self._lib.cap_courant
        rc = fn(self._h, self._f(f1), self._f(f2), self._f(f3), self._f(vol), self._f(su),
                C.byref(ts), C.byref(tm), 1 if conumfix else 0, float(conumfixvalue), out)
        self._check(rc, fn.__name__, allow=(CAP_OK,))
        return out[0], out[1], ts.value, tm.value

    def fluxmc(self, f, d, fmcor):
        st = self._fields(f)
        rc = self._lib.cap_fluxmc(self._h, C.byref(st), *self._dir(d),
                                  self._f(f[("fx", "fy", "fz")[d]]), self._f(fmcor))
        self._check(rc, "cap_fluxmc", allow=(CAP_OK,))

    def continuity(self, f1, f2, f3):
        out = (C.c_double * 2)()
        rc = self._lib.cap_continuity(self._h, self._f(f1), self._f(f2), self._f(f3), out)
        self._check(rc, "cap_continuity", allow=(CAP_OK,))
        return list(out)

    # ---------------------------------------------- pressure-correction drivers
    def _flow(self, f):
        """cap_flow from a dict of device tensors (missing names: NULL)."""
        return _Flow(*[self._f(f.get(n), n, 3 if n.startswith("grad") else 1)
                       for n in FLOW_NAMES])

    @staticmethod
    def _reports(reps, n):
        return [dict(resor=r.resor, resl=r.resl, iters=r.iters, status=r.status, ms=r.ms,
                     converged=r.status == CAP_OK) for r in reps[:n]]

    def calcp(self, f, opts):
        """`call calcp` on device-resident module state f (dict of tensors);
        returns (solver reports, continuity sums per corrector)."""
        st = self._flow(f)
        n = opts.ncorr * opts.npcor
        reps = (_Report * n)()
        cont = (C.c_double * (3 * opts.ncorr))()
        rc = self._lib.cap_calcp(self._h, C.byref(st), C.byref(opts), reps, cont)
        self._check(rc, "cap_calcp")
        return self._reports(reps, n), [tuple(cont[3 * q:3 * q + 3]) for q in range(opts.ncorr)]

    def piso(self, f, opts, pimple=False):
        """`call PISO_multiple_correction` (or PIMPLE_multiple_correction)."""
        st = self._flow(f)
        n = opts.ncorr * opts.npcor
        reps = (_Report * n)()
        cont = (C.c_double * (3 * opts.ncorr))()
        rc = self._lib.cap_piso(self._h, C.byref(st), C.byref(opts), 1 if pimple else 0, reps, cont)
        self._check(rc, "cap_piso")
        return self._reports(reps, n), [tuple(cont[3 * q:3 * q + 3]) for q in range(opts.ncorr)]

    def piso_guess(self, f, opts, prev, net=None, pimple=False):
        """piso with the last corrector's p solve started from the learned
        guess (net: a UNet on this context) or from prev, the previous time
        step's p, on the solve range (net None), when that start lowers the
        L1 residual of the reference's start; the solve stops at the absolute
        residual the reference's start stops at.  Returns (reports,
        continuity, accepted, (L1 of the reference's start, of the start
        used))."""
        st = self._flow(f)
        n = opts.ncorr * opts.npcor
        reps = (_Report * n)()
        cont = (C.c_double * (3 * opts.ncorr))()
        acc = C.c_int()
        l1 = (C.c_double * 2)()
        rc = self._lib.cap_piso_guess(self._h, C.byref(st), C.byref(opts), 1 if pimple else 0,
                                      net._h if net is not None else None, self._f(prev, "prev"),
                                      reps, cont, C.byref(acc), l1)
        self._check(rc, "cap_piso_guess")
        return (self._reports(reps, n), [tuple(cont[3 * q:3 * q + 3]) for q in range(opts.ncorr)],
                bool(acc.value), (l1[0], l1[1]))

    def outbc(self, f, opts):
        st = self._flow(f)
        self._check(self._lib.cap_outbc(self._h, C.byref(st), C.byref(opts)), "cap_outbc",
                    allow=(CAP_OK,))

    def correct_bc(self, f):
        st = self._flow(f)
        self._check(self._lib.cap_correct_bc(self._h, C.byref(st)), "cap_correct_bc",
                    allow=(CAP_OK,))

    def gethbya(self, f, opts):
        st = self._flow(f)
        self._check(self._lib.cap_gethbya(self._h, C.byref(st), C.byref(opts)), "cap_gethbya",
                    allow=(CAP_OK,))

    def calcuvw(self, f, spv, opts, uvw):
        """`call calcuvw`; returns the three sipsol reports (u, v, w)."""
        st = self._flow(f)
        reps = (_Report * 3)()
        rc = self._lib.cap_calcuvw(self._h, C.byref(st), self._f(spv), C.byref(opts), C.byref(uvw), reps)
        self._check(rc, "cap_calcuvw")
        return self._reports(reps, 3)

    def find_strain_rate(self, f, strain, vorticity):
        st = self._flow(f)
        rc = self._lib.cap_find_strain_rate(self._h, C.byref(st), self._f(strain), self._f(vorticity))
        self._check(rc, "cap_find_strain_rate", allow=(CAP_OK,))

    def calc_vis_les(self, f, strain, walldist, les):
        st = self._flow(f)
        rc = self._lib.cap_calc_vis_les(self._h, C.byref(st), self._f(strain), self._f(walldist),
                                        C.byref(les))
        self._check(rc, "cap_calc_vis_les", allow=(CAP_OK,))

    # --- grid geometry -----------------------------------------------------
    AR_FIELDS = ("ar1x", "ar1y", "ar1z", "ar2x", "ar2y", "ar2z", "ar3x", "ar3y", "ar3z")

    def face_areas(self, g):
        """init.f90:49-144: g's ar1x .. ar3z from its nodes x, y, z (in place)."""
        A = (C.c_void_p * 9)(*[self._f(g[q], q) for q in self.AR_FIELDS])
        rc = self._lib.cap_face_areas(self._h, self._f(g["x"], "x"), self._f(g["y"], "y"),
                                      self._f(g["z"], "z"), A)
        self._check(rc, "cap_face_areas", allow=(CAP_OK,))

    def intfac(self, g):
        """init.f90:147-149: g's fx, fy, fz from its nodes and centres (in place)."""
        G = (C.c_void_p * 6)(*[self._f(g[q], q) for q in ("x", "y", "z", "xc", "yc", "zc")])
        rc = self._lib.cap_intfac(self._h, G, self._f(g["fx"], "fx"), self._f(g["fy"], "fy"),
                                  self._f(g["fz"], "fz"))
        self._check(rc, "cap_intfac", allow=(CAP_OK,))

    def build_lognormal(self, coef, seed, sigma=1.0):
        st = coef._struct(self)
        rc = self._lib.cap_build_lognormal(self._h, seed, sigma, C.byref(st))
        self._check(rc, "cap_build_lognormal", allow=(CAP_OK,))


GRID_ARRAYS = ("x", "y", "z", "xc", "yc", "zc", "vol", "fx", "fy", "fz",
               "ar1x", "ar1y", "ar1z", "ar2x", "ar2y", "ar2z", "ar3x", "ar3y", "ar3z")
GRID_LB = ("lbw", "lbe", "lbs", "lbn", "lbb", "lbt")
GRID_FLAGS = ("low", "loe", "los", "lon", "lob", "lot", "lout", "lthermbc", "lconbc")


def read_grid(path):
    """The reference's grid file (read_grid.f90:126-161) through
    cap_grid_dims / cap_read_grid: returns (dims, arrays, lb, flags) with
    numpy arrays named as the reference's module variables."""
    import numpy as np
    L = load()
    ni, nj, nk = C.c_int(), C.c_int(), C.c_int()
    rc = L.cap_grid_dims(os.fsencode(path), C.byref(ni), C.byref(nj), C.byref(nk))
    if rc != CAP_OK:
        raise CapError(f"cap_grid_dims: {path} is not a readable grid file (status {rc})")
    dims = (ni.value, nj.value, nk.value)
    n = dims[0] * dims[1] * dims[2]
    arrays = {q: np.zeros(n) for q in GRID_ARRAYS}
    lens = (dims[1] * dims[2],) * 2 + (dims[0] * dims[2],) * 2 + (dims[0] * dims[1],) * 2
    lb = {q: np.zeros(m, np.int32) for q, m in zip(GRID_LB, lens)}
    flags = (C.c_int * 9)()
    A = (C.c_void_p * 19)(*[arrays[q].ctypes.data for q in GRID_ARRAYS])
    B = (C.c_void_p * 6)(*[lb[q].ctypes.data for q in GRID_LB])
    rc = L.cap_read_grid(os.fsencode(path), *dims, A, B, flags)
    if rc != CAP_OK:
        raise CapError(f"cap_read_grid: {path} is malformed (status {rc})")
    return dims, arrays, lb, dict(zip(GRID_FLAGS, (bool(v) for v in flags)))


def _fptr(a, n=None, device=None, what="array"):
    """float32 array (numpy, or torch on the device)"""
    return _check_array(a, "float32", n, device, what)


class UNet:
    """The learned initial guess (csrc/cap_unet.hip) on one context's grid:
    nx x ny x nz interior cells (x = the reference's j index, fastest)."""

    def __init__(self, ctx, nx, ny, nz, seed=1):
        self._ctx = ctx
        self._lib = ctx._lib
        self.shape = (nz, ny, nx)
        h = C.c_void_p()
        rc = self._lib.cap_unet_create(ctx._h, nx, ny, nz, seed, C.byref(h))
        ctx._check(rc, "cap_unet_create", allow=(CAP_OK,))
        self._h = h
        n = C.c_long()
        self._lib.cap_unet_num_params(self._h, C.byref(n))
        self.num_params = n.value

    def close(self):
        if self._h:
            self._lib.cap_unet_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _in(self, x):
        nz, ny, nx = self.shape
        return _fptr(x, 3 * nx * ny * nz, self._ctx.device_pointers, "unet input")

    def _out(self, y):
        nz, ny, nx = self.shape
        return _fptr(y, nx * ny * nz, self._ctx.device_pointers, "unet output")

    def _grads(self, g):
        if g is None:
            return None
        return _fptr(g, self.num_params, self._ctx.device_pointers, "unet gradients")

    def _chk(self, rc, what):
        self._ctx._check(rc, what, allow=(CAP_OK,))

    def get_params(self):
        import numpy as np
        p = np.zeros(self.num_params, np.float32)
        self._chk(self._lib.cap_unet_get_params(self._h, p.ctypes.data), "cap_unet_get_params")
        return p

    def set_params(self, p):
        import numpy as np
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._chk(self._lib.cap_unet_set_params(self._h, p.ctypes.data), "cap_unet_set_params")

    def forward(self, x, out):
        self._chk(self._lib.cap_unet_forward(self._h, self._in(x), self._out(out)), "cap_unet_forward")

    def gradients(self, x, target, grads=None):
        loss = C.c_float()
        rc = self._lib.cap_unet_gradients(self._h, self._in(x), self._out(target), C.byref(loss),
                                          self._grads(grads))
        self._chk(rc, "cap_unet_gradients")
        return loss.value

    def train_step(self, x, target, lr):
        loss = C.c_float()
        rc = self._lib.cap_unet_train_step(self._h, self._in(x), self._out(target), float(lr),
                                           C.byref(loss))
        self._chk(rc, "cap_unet_train_step")
        return loss.value

    def guess(self, fi, coef, prev=None):
        """fi (device, reference layout) becomes the guess; returns
        (s, a, ps, alpha, beta), the normalisers and the fitted weights."""
        sc = (C.c_double * 5)()
        st = coef._struct(self._ctx)
        rc = self._lib.cap_unet_guess(self._ctx._h, self._h, self._ctx._f(fi, 'fi'),
                                      self._ctx._f(prev, 'prev'), C.byref(st), sc)
        self._chk(rc, "cap_unet_guess")
        return tuple(sc)

    def pcgsip(self, fi, coef, prev=None, alfa=0.0, sor=1e-8, nsw=2000):
        rep = _Report()
        acc = C.c_int()
        st = coef._struct(self._ctx)
        rc = self._lib.cap_pcgsip_guess(self._ctx._h, self._h, self._ctx._f(fi, 'fi'),
                                        self._ctx._f(prev, 'prev'), C.byref(st),
                                        alfa, sor, nsw, C.byref(rep), C.byref(acc))
        self._ctx._check(rc, "cap_pcgsip_guess",
                         allow=(CAP_OK, CAP_MAXITER, CAP_GUESS_REJECTED))
        return dict(resor=rep.resor, resl=rep.resl, iters=rep.iters, status=rep.status,
                    ms=rep.ms, converged=rep.status == CAP_OK, accepted=bool(acc.value))
