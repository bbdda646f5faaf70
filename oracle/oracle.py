"""CPU oracle for the p' hot path -- TEST INFRASTRUCTURE ONLY.

ctypes front-end of ``liboracle.so`` (the C restatement in ``cap_oracle.c``)
plus a runner for the reference-compiled driver in ``oracle/_ref/`` (build
container only).  Only tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg import this module, and only as the checker.

Arrays are numpy float64 vectors of length ni*nj*nk in the reference's padded
j-fastest layout (``ijk = (k-1)*ni*nj + (i-1)*nj + j``, C offset ``ijk-1``);
``as_grid(a, dims)`` views one as ``[k, i, j]`` (0-based).
"""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")
# the same driver linked with update_values_at_ghostcells.f90 (periodic variant)
REF_DRIVER_PERIODIC = os.path.join(HERE, "_ref", "ref_driver_periodic")
DROPIN_DRIVER = os.path.join(os.path.dirname(HERE), "tests", "fortran", "_build", "dropin_driver")

SYS_FIELDS = ("ae", "aw", "an", "as_", "at", "ab", "ap", "su", "res",
              "bb", "bs", "bw", "bp", "bn", "be", "bt")

SMALL = float(np.float32(1e-30))
ALFA = 0.92


class _Sys(C.Structure):
    _fields_ = [(f, C.POINTER(C.c_double)) for f in SYS_FIELDS]


class _Report(C.Structure):
    _fields_ = [("resor", C.c_double), ("resl", C.c_double),
                ("iters", C.c_int), ("converged", C.c_int)]


# module state of the drivers, in the order of orc_flow
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
    """orc_drv_opts: the module settings the drivers read."""
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
    """orc_uvw_opts: calcuvw's settings beyond the driver settings."""
    _fields_ = [("schemes", C.c_int), ("gds", C.c_double), ("urfr", C.c_double * 3),
                ("urfm", C.c_double * 3), ("sor", C.c_double * 3), ("nsw", C.c_int * 3),
                ("gradpcmf", C.c_double), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("top_k", C.c_int)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        i, u64, d = C.c_int, C.c_uint64, C.c_double
        L.orc_u01.argtypes = [u64, u64, u64]
        L.orc_u01.restype = d
        L.orc_normal.argtypes = [u64, u64, u64]
        L.orc_normal.restype = d
        L.orc_hash.argtypes = [u64, u64, u64]
        L.orc_hash.restype = u64
        for fn in ("orc_gen_log", "orc_gen_cos2pi", "orc_gen_exp"):
            getattr(L, fn).argtypes = [d]
            getattr(L, fn).restype = d
        L.orc_build_lognormal.argtypes = [i, i, i, u64, d, C.POINTER(_Sys)]
        L.orc_build_general.argtypes = [i, i, i, u64, C.POINTER(_Sys), dp]
        for fn in ("orc_sipsol", "orc_pcgsip", "orc_cgstab_sip"):
            getattr(L, fn).argtypes = [i, i, i, dp, C.POINTER(_Sys), d, d, i,
                                       C.POINTER(_Report)]
        L.orc_iccg.argtypes = [i, i, i, dp, C.POINTER(_Sys), d, i, C.POINTER(_Report)]
        L.orc_cgstab.argtypes = [i, i, i, dp, C.POINTER(_Sys), d, i, C.POINTER(_Report)]
        L.orc_pcg_jacobi.argtypes = [i, i, i, dp, C.POINTER(_Sys), d, i, C.POINTER(_Report)]
        L.orc_sip_factor.argtypes = [i, i, i, i, i, C.POINTER(_Sys), d]
        L.orc_sip_apply.argtypes = [i, i, i, i, i, C.POINTER(_Sys), dp, dp]
        L.orc_spmv.argtypes = [i, i, i, i, i, C.POINTER(_Sys), dp, dp]
        L.orc_grad_gauss.argtypes = [i, i, i] + [dp] * 17
        L.orc_bpres.argtypes = [i, i, i, dp, dp, dp, dp]
        L.orc_pcgsip_time_iters.argtypes = [i, i, i, dp, C.POINTER(_Sys), d, i]
        L.orc_pcgsip_time_iters.restype = d
        L.orc_set_plane_sums.argtypes = [i]
        Fl, Op = C.POINTER(_Flow), C.POINTER(DrvOpts)
        L.orc_calcp.argtypes = [i, i, i, Fl, Op, C.POINTER(_Report), dp]
        L.orc_piso.argtypes = [i, i, i, Fl, Op, i, C.POINTER(_Report), dp]
        L.orc_outbc.argtypes = [i, i, i, Fl, Op]
        L.orc_correct_bc.argtypes = [i, i, i, Fl]
        L.orc_gethbya.argtypes = [i, i, i, Fl, Op]
        L.orc_strain_rate.argtypes = [i, i, i, dp, dp, dp, dp, dp]
        L.orc_calc_vis_les.argtypes = [i, i, i] + [dp] * 15 + [d, d, d, d, d, i, i, i, i]
        L.orc_calcuvw.argtypes = [i, i, i, Fl, dp, Op, C.POINTER(UvwOpts), C.POINTER(_Report)]
        L.orc_face_areas.argtypes = [i, i, i, dp, dp, dp, C.POINTER(dp)]
        ip = C.POINTER(C.c_int)
        L.orc_ghostcells.argtypes = [i, i, i, C.POINTER(dp), C.POINTER(dp)] + [ip] * 6 + [dp]
        L.orc_ghostcells_periodic.argtypes = [i, i, i, C.POINTER(dp), dp]
        L.orc_find_intersection_point.argtypes = [dp] * 6
        L.orc_intfac.argtypes = [i, i, i, i, i, i] + [dp] * 7
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


class System:
    """The module arrays `coef`/`coefb` that the reference solvers use."""

    def __init__(self, dims):
        self.dims = tuple(int(x) for x in dims)
        n = self.dims[0] * self.dims[1] * self.dims[2]
        for f in SYS_FIELDS:
            setattr(self, f, np.zeros(n))
        self.ap[:] = 1.0

    def copy(self):
        s = System(self.dims)
        for f in SYS_FIELDS:
            getattr(s, f)[:] = getattr(self, f)
        return s

    def _struct(self):
        return _Sys(*[_p(getattr(self, f)) for f in SYS_FIELDS])

    @property
    def n(self):
        return self.dims[0] * self.dims[1] * self.dims[2]


def as_grid(a, dims):
    ni, nj, nk = dims
    return a.reshape(nk, ni, nj)


def build_lognormal(dims, seed, sigma=1.0):
    s = System(dims)
    lib().orc_build_lognormal(*s.dims, seed, sigma, C.byref(s._struct()))
    return s


def build_general(dims, seed):
    s = System(dims)
    fi = np.zeros(s.n)
    lib().orc_build_general(*s.dims, seed, C.byref(s._struct()), _p(fi))
    return s, fi


def _solve(fn, s, fi, *args):
    rep = _Report()
    st = s._struct()
    fn(*s.dims, _p(fi), C.byref(st), *args, C.byref(rep))
    return dict(resor=rep.resor, resl=rep.resl, iters=rep.iters,
                converged=bool(rep.converged))


def sipsol(s, fi, alfa=ALFA, sor=1e-8, nsw=1000):
    return _solve(lib().orc_sipsol, s, fi, alfa, sor, nsw)


def pcgsip(s, fi, alfa=ALFA, sor=1e-8, nsw=1000):
    return _solve(lib().orc_pcgsip, s, fi, alfa, sor, nsw)


def cgstab_sip(s, fi, alfa=ALFA, sor=1e-8, nsw=1000):
    return _solve(lib().orc_cgstab_sip, s, fi, alfa, sor, nsw)


def iccg(s, fi, sor=1e-8, nsw=1000):
    return _solve(lib().orc_iccg, s, fi, sor, nsw)


def cgstab(s, fi, sor=1e-8, nsw=1000):
    """cgstab.f90: BiCGStab with the diagonal ILU(0) preconditioner"""
    return _solve(lib().orc_cgstab, s, fi, sor, nsw)


def pcg_jacobi(s, fi, sor=1e-8, nsw=1000):
    """pcg-jacobi.f90: CG with the Jacobi preconditioner"""
    return _solve(lib().orc_pcg_jacobi, s, fi, sor, nsw)


class plane_sums:
    """Context manager: the oracle's scalar products and norms are summed per
    k-plane (mode 1) or as compensated (TwoSum) running sums, i.e. nearly
    exactly (mode 2, the GPU's arithmetic), instead of in one plain running
    sum (mode 0, the reference's)."""

    def __init__(self, mode=1):
        self.mode = mode

    def __enter__(self):
        lib().orc_set_plane_sums(self.mode)

    def __exit__(self, *exc):
        lib().orc_set_plane_sums(0)


def order_sensitivity(solver, s, fi, **kw):
    """max|fi_a - fi_b| / max|fi_a| between the reference's summation and the
    per-plane and compensated summations (the larger of the two): how far the
    solve moves when only the summation order / rounding changes."""
    s1, f1 = s.copy(), fi.copy()
    solver(s1, f1, **kw)
    worst = 0.0
    for mode in (1, 2):
        s2, f2 = s.copy(), fi.copy()
        with plane_sums(mode):
            solver(s2, f2, **kw)
        worst = max(worst, float(np.max(np.abs(f1 - f2)) / max(np.max(np.abs(f1)), 1e-300)))
    return worst


def accurate_sums(solver, s, fi, **kw):
    """The solve with compensated (nearly exact) scalar products and norms --
    the arithmetic of the GPU's reductions.  Returns (report, s, fi)."""
    s2, f2 = s.copy(), fi.copy()
    with plane_sums(2):
        rep = solver(s2, f2, **kw)
    return rep, s2, f2


def sip_factor(s, k_lo, k_hi, alfa=ALFA):
    lib().orc_sip_factor(*s.dims, k_lo, k_hi, C.byref(s._struct()), alfa)


def sip_apply(s, k_lo, k_hi, rhs):
    z = np.zeros(s.n)
    lib().orc_sip_apply(*s.dims, k_lo, k_hi, C.byref(s._struct()), _p(rhs), _p(z))
    return z


def spmv(s, k_lo, k_hi, p):
    z = np.zeros(s.n)
    lib().orc_spmv(*s.dims, k_lo, k_hi, C.byref(s._struct()), _p(p), _p(z))
    return z


def pcgsip_time_iters(s, fi, n_iters, alfa=ALFA):
    return lib().orc_pcgsip_time_iters(*s.dims, _p(fi), C.byref(s._struct()), alfa, n_iters)


GEOM_FIELDS = ("fx", "fy", "fz", "ar1x", "ar1y", "ar1z", "ar2x", "ar2y", "ar2z",
               "ar3x", "ar3y", "ar3z", "vol")


def grad_gauss(dims, u, geom):
    n = dims[0] * dims[1] * dims[2]
    out = [np.zeros(n) for _ in range(3)]
    lib().orc_grad_gauss(*dims, _p(u), *[_p(geom[f]) for f in GEOM_FIELDS],
                         *[_p(o) for o in out])
    return out


def bpres(dims, p, geom):
    lib().orc_bpres(*dims, _p(p), _p(geom["fx"]), _p(geom["fy"]), _p(geom["fz"]))


# the oracle's orc_fields: the arrays the p' assembly and fluxuvw read
FIELD_NAMES = ("x", "y", "z", "xc", "yc", "zc", "vol", "fx", "fy", "fz", "ar1x", "ar1y",
               "ar1z", "ar2x", "ar2y", "ar2z", "ar3x", "ar3y", "ar3z", "den", "apu", "apv",
               "apw", "u", "v", "w", "p", "gradu", "gradv", "gradw", "gradp")


class _Fields(C.Structure):
    _fields_ = [(q, C.POINTER(C.c_double)) for q in FIELD_NAMES]


def fluxuvw(dims, f, vis, d, out, flags, gds):
    """fluxuvw-overrelaxed-correction.f90 for direction d (0 east, 1 north,
    2 top) on the fields f (FIELD_NAMES arrays) with fcf = f1 / f2 / f3 of f;
    out: dict of acfe, acfw, su, sv, sw, bp, bt, bb, updated in place."""
    ni, nj, nk = dims
    nij = ni * nj
    idew, idns, idtb = [(nj, 1, nij), (1, nij, nj), (nij, nj, 1)][d]
    dp = C.POINTER(C.c_double)
    fs = _Fields(*[_p(f[q]) for q in FIELD_NAMES])
    L = lib()
    L.orc_fluxuvw.argtypes = [C.c_int] * 3 + [C.c_void_p, dp] + [C.c_int] * 3 + [dp] * 13 + \
        [C.c_int, C.c_double]
    arv = [f[f"ar{d + 1}{c}"] for c in "xyz"]
    L.orc_fluxuvw(*dims, C.byref(fs), _p(vis), idew, idns, idtb, *map(_p, arv),
                  _p(f[("fx", "fy", "fz")[d]]), _p(out["acfe"]), _p(out["acfw"]),
                  _p(f[("f1", "f2", "f3")[d]]), _p(out["su"]), _p(out["sv"]), _p(out["sw"]),
                  _p(out["bp"]), _p(out["bt"]), _p(out["bb"]), int(flags), float(gds))


def ghostcells(dims, var, geo, lb, inl):
    """update_values_at_ghostcells_nonperiodic.f90 in place on var = [u, v,
    w, te, ed, t, vis]; geo = [fx, fy, fz, ar1x .. ar3z]; lb = int32 [lbw,
    lbe, lbs, lbn, lbb, lbt]; inl = (uin, vin, win, tein, edin, tin)."""
    dp = C.POINTER(C.c_double)
    V = (dp * 7)(*[_p(a) for a in var])
    G = (dp * 12)(*[_p(a) for a in geo])
    ip = C.POINTER(C.c_int)
    for a in lb:
        assert a.dtype == np.int32 and a.flags.c_contiguous
    inl_ = np.ascontiguousarray(inl, dtype=np.float64)
    lib().orc_ghostcells(*dims, V, G, *[a.ctypes.data_as(ip) for a in lb], _p(inl_))


def ghostcells_periodic(dims, var, fz):
    """update_values_at_ghostcells.f90 (periodic i, j; walls at k) in place on
    var = [u, v, w, te, ed, t, vis] (t untouched)."""
    dp = C.POINTER(C.c_double)
    V = (dp * 7)(*[_p(a) for a in var])
    lib().orc_ghostcells_periodic(*dims, V, _p(fz))


AR_FIELDS = ("ar1x", "ar1y", "ar1z", "ar2x", "ar2y", "ar2z", "ar3x", "ar3y", "ar3z")


def face_areas(dims, x, y, z, ar):
    """init.f90:49-144 into ar (dict of the nine ar* arrays, updated in place
    on the reference's loop ranges)"""
    arr = (C.POINTER(C.c_double) * 9)(*[_p(ar[f]) for f in AR_FIELDS])
    lib().orc_face_areas(*dims, _p(x), _p(y), _p(z), arr)


def find_intersection_point(p1, p2, p3, p4, p5):
    pts = [np.ascontiguousarray(q, dtype=np.float64) for q in (p1, p2, p3, p4, p5)]
    out = np.zeros(3)
    lib().orc_find_intersection_point(*[_p(q) for q in pts], _p(out))
    return out


def intfac(dims, g, lam=None):
    """intfac.f90 for the three directions as init.f90:147-149 calls it:
    returns (fx, fy, fz), starting from g's fx, fy, fz (or lam) outside the
    loop range"""
    ni, nj, nk = dims
    nij = ni * nj
    out = []
    for q, (de, dn, dt) in zip(("fx", "fy", "fz"), ((nj, 1, nij), (1, nj, nij), (nij, 1, nj))):
        a = (lam[q] if lam is not None else g[q]).copy()
        lib().orc_intfac(*dims, de, dn, dt, *[_p(g[c]) for c in ("x", "y", "z", "xc", "yc", "zc")],
                         _p(a))
        out.append(a)
    return out


def rng_u01(seed, stream, idx):
    return lib().orc_u01(seed, stream, idx)


def rng_normal(seed, stream, idx):
    return lib().orc_normal(seed, stream, idx)


def gen_log(x):
    """the generator's libm-free natural log"""
    return lib().orc_gen_log(x)


def gen_cos2pi(u):
    """the generator's libm-free cos(2 pi u), u in [0, 1)"""
    return lib().orc_gen_cos2pi(u)


def gen_exp(x):
    """the generator's libm-free exp"""
    return lib().orc_gen_exp(x)


# ------------------------------------------------- pressure-correction drivers --
def _flow_struct(f):
    ptrs = []
    for name in FLOW_NAMES:
        a = f.get(name)
        if a is None:
            ptrs.append(None)
        else:
            assert a.dtype == np.float64 and a.flags.c_contiguous, name
            ptrs.append(a.ctypes.data)
    return _Flow(*ptrs)


def _reports(reps, n):
    return [dict(resor=r.resor, resl=r.resl, iters=r.iters, converged=bool(r.converged))
            for r in reps[:n]]


def calcp(dims, f, opts):
    """calcp-multiple_correction_SIMPLE on the dict of arrays f (modified in
    place); returns (solver reports, continuity sums per corrector)."""
    st = _flow_struct(f)
    n = opts.ncorr * opts.npcor
    reps = (_Report * n)()
    cont = np.zeros(3 * opts.ncorr)
    lib().orc_calcp(*dims, C.byref(st), C.byref(opts), reps, _p(cont))
    return _reports(reps, n), [tuple(cont[3 * q:3 * q + 3]) for q in range(opts.ncorr)]


def piso(dims, f, opts, pimple=False):
    st = _flow_struct(f)
    n = opts.ncorr * opts.npcor
    reps = (_Report * n)()
    cont = np.zeros(3 * opts.ncorr)
    lib().orc_piso(*dims, C.byref(st), C.byref(opts), 1 if pimple else 0, reps, _p(cont))
    return _reports(reps, n), [tuple(cont[3 * q:3 * q + 3]) for q in range(opts.ncorr)]


def outbc(dims, f, opts):
    st = _flow_struct(f)
    lib().orc_outbc(*dims, C.byref(st), C.byref(opts))


def correct_bc(dims, f):
    st = _flow_struct(f)
    lib().orc_correct_bc(*dims, C.byref(st))


def gethbya(dims, f, opts):
    st = _flow_struct(f)
    lib().orc_gethbya(*dims, C.byref(st), C.byref(opts))


def strain_rate(dims, gradu, gradv, gradw):
    n = dims[0] * dims[1] * dims[2]
    strain, vort = np.zeros(n), np.zeros(n)
    lib().orc_strain_rate(*dims, _p(gradu), _p(gradv), _p(gradw), _p(strain), _p(vort))
    return strain, vort


LES_IN = ("u", "v", "w", "vol", "den", "fz", "xc", "yc", "zc", "ar3x", "ar3y", "ar3z")


def calc_vis_les(dims, f, vis, par):
    """calc_vis_les on the dict f (LES_IN arrays plus strain, walldist); vis
    is updated in place; par: viscos, densit, cappa, elog, urf, nx, ny, nz"""
    lib().orc_calc_vis_les(*dims, *[_p(f[q]) for q in LES_IN], _p(f["strain"]),
                           _p(f["walldist"]), _p(vis), par["viscos"], par["densit"],
                           par["cappa"], par["elog"], par["urf"], par["nx"], par["ny"], par["nz"],
                           par.get("top_k", 0))


def calcuvw(dims, f, spv, opts, uvw):
    st = _flow_struct(f)
    reps = (_Report * 3)()
    lib().orc_calcuvw(*dims, C.byref(st), _p(spv), C.byref(opts), C.byref(uvw), reps)
    return _reports(reps, 3)


# ----------------------------------------------------- reference driver --
def ref_available():
    return os.path.exists(REF_DRIVER)


_LOG_RE = re.compile(r"Final residual =\s*([-+0-9.Ee]+), No Iterations (\d+)")


def run_ref(mode, dims, arrays, nsw=0, alfa=ALFA, sor=1e-8, driver=REF_DRIVER):
    """Run the reference-compiled driver (or, with driver=DROPIN_DRIVER, the
    host program linked with the Fortran drop-ins); returns
    (outputs, resor, log)."""
    ni, nj, nk = dims
    n = ni * nj * nk
    with tempfile.TemporaryDirectory() as td:
        fin, fout, flog = (os.path.join(td, x) for x in ("in.bin", "out.bin", "log.txt"))
        with open(fin, "wb") as f:
            np.array([ni, nj, nk, nsw], dtype=np.int32).tofile(f)
            np.array([alfa, sor], dtype=np.float64).tofile(f)
            for a in arrays:
                dt = np.int32 if getattr(a, "dtype", None) == np.int32 else np.float64
                np.ascontiguousarray(a, dtype=dt).tofile(f)
        subprocess.run([driver, mode, fin, fout, flog], check=True,
                       capture_output=True)
        raw = np.fromfile(fout, dtype=np.float64)
        log = open(flog).read()
    if mode in ("continuity", "fip"):
        return [raw.copy()], None, log
    nout = {"sipsol": 9, "cgstab_sip": 9, "pcgsip": 9, "iccg": 9, "cgstab": 9, "pcgjac": 9, "grad_gauss": 3,
            "bpres": 1, "fluxmass": 5, "piso_assemble": 5, "assemble_p": 5, "fluxmc": 5,
            "fluxuvw": 8, "ghost": 7, "outbc": 15, "cbc": 5, "les": 1, "intfac": 3}[mode]
    outs = [raw[i * n:(i + 1) * n].copy() for i in range(nout)]
    resor = float(raw[nout * n]) if raw.size > nout * n else None
    return outs, resor, log


def parse_log(log):
    m = _LOG_RE.search(log)
    if not m:
        return None
    return dict(resl=float(m.group(1)), iters=int(m.group(2)))
